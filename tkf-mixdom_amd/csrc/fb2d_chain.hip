// tkfmi: chained cell-per-lane 2D pair-HMM forward/backward + expected
// counts for gfx950 (MI355X), TKF91/TKF92 (S,M,I,D,E) and MixDom
// (2 + 5KF states, K*F <= 3).  Reference algorithm:
// python/tkfmixdom/jax/dp/hmm.py (forward hmm.py:1186-1292, backward
// hmm.py:1295-1437, posteriors and counts hmm.py:1930-1971).
//
// The per-cell arithmetic is the one of the single-workgroup kernel in
// fb2d_v2.hip (one lattice cell per lane, probability space with a binary
// exponent per cell, DPP row broadcasts of P, MFMA counts).  What differs is
// how a pair is spread over the machine:
//  * A pair's bands of 64 rows are cut into a chain of G workgroups of at
//    most W waves, one band per wave, so a long pair runs on G compute units
//    at one wave per SIMD instead of on one compute unit at two.
//  * Inside a workgroup, waves run 64 steps apart in lock-step and pass the
//    boundary row through LDS (as in the single-workgroup kernel).  Between
//    workgroups of a chain the boundary row goes through global memory as
//    device-scope atomic stores; every CQ columns the producer waits for
//    them to complete and stores a progress counter, the consumer acquires
//    it and copies the next CQ entries into LDS.
//  * Workgroups take tickets from a counter in launch order.  Tickets go by
//    critical path (the steps from a workgroup's start to the end of its
//    pair's chain), which falls along a chain, so a pair's chain is ticketed
//    in dependency order (top band first in the forward, bottom band first
//    in the backward).  A workgroup therefore waits only on workgroups that
//    already hold a ticket, i.e. are running: no deadlock whatever the
//    occupancy.  Every wait is bounded; on a time-out the launch raises an
//    abort flag, all waits end, and the affected pairs get a NaN
//    log-likelihood.
//  * A backward workgroup has one more wave, the count wave: the recursion
//    waves stage their cells' count operands in LDS each step and the count
//    wave runs the expected-count MFMAs on its own SIMD.
//  * The forward and the backward are separate launches on the class's
//    stream; log P and its binary exponent pass through a per-pair slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fb2d_cell.h"
#include "fb2d_chain.h"
#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_v2 {

// status[0] = 1 if the call's abort flag was raised (after every class joined).
__global__ void raise_status_kernel(const int32_t* __restrict__ abortf, int32_t* __restrict__ status) {
  if (threadIdx.x == 0 && __hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    status[0] = 1;
}

// ============================================================== forward
template <class L, int W>
__global__ __launch_bounds__(W * 64) void fbc_fwd_kernel(CParams Pr, const double* __restrict__ pmat,
                                                         const CDesc* __restrict__ descs) {
  constexpr int NB = L::NB, NG = L::NG;
  constexpr int nM = count_type<L, ST_M>();
  constexpr int nD = count_type<L, ST_D>();
  using LY = CLds<L, W>;
  constexpr int NX = LY::NX;
  constexpr int EIDX = L::EIDX;
  static_assert(NB <= SLOTS - 1, "exponent slot");
  extern __shared__ double smem[];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const CDesc pd = descs[take_ticket(Pr.ticket, Pr.abortf)];
  const int Lx = pd.Lx, Ly = pd.Ly, NW = pd.nw;
  const int S = Pr.S;
  const int BSTRIDE = Ly + 64;

  double* s_tab = smem + LY::tab;
  double* s_ring = smem + LY::ring(false);
  double* s_xc = smem + LY::xc(false);
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + LY::seq_bytes(false);
  uint8_t* s_y = s_x + LY::xbytes;
  stage_inputs<L, W>(Pr, pd, pmat + pd.pm_off, s_tab, s_x, s_y);

  const double* __restrict__ pm = pmat + pd.pm_off;
  const double* p_start = smem + LY::misc;
  const double* p_end = p_start + NB;
  double prow[NB];  // prow[s] in lane l: P[s][l % 16] (0 past NB)
#pragma unroll
  for (int s = 0; s < NB; ++s) prow[s] = valu_copy((lane & 15) < NB ? pm[s * NB + (lane & 15)] : 0.0);

  const bool want_lat = !(Pr.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lat && !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* lat = Pr.lattice + pd.lat_off;
  const bool has_band = w < NW;
  const int band = pd.band0 + w;
  const int i_band = 64 * band + lane;
  const bool xin = (w == 0) && pd.fin >= 0;
  const bool xout = (w == NW - 1) && pd.fout >= 0;
  bool ok_run = true;
  __syncthreads();

  double hp[NB];
  double rM_last[nM];
#pragma unroll
  for (int b = 0; b < NB; ++b) hp[b] = 0.0;
#pragma unroll
  for (int m = 0; m < nM; ++m) rM_last[m] = 0.0;
  int eH = EXP_NONE, eM_last = EXP_NONE;
  // The lane's row, hence x and its D emissions, is fixed; the column's y
  // moves one lane down per step with the cells (DPP), lane 0 reading the
  // next one from LDS a step ahead.  One LDS round trip per step remains.
  const int xi = s_x[64 * w + lane];
  double emD[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) emD[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + A_PAD + xi];
  int yj_prev = 0;
  int y0_next = s_y[0];
  const int T = pd.T;
  // A wave's band is in range for steps 64 w .. 64 w + 63 + Ly; outside
  // them every lane is invalid and the wave only keeps the barrier (its
  // carried values are zero there, as the skipped steps would leave them).
  const int t_on = 64 * w, t_off = 64 * w + 63 + Ly;
  for (int t = 0; t < T; ++t) {
    if (has_band && t >= t_on && t <= t_off) {
      int i = i_band;
      asm volatile("" : "+v"(i));  // keeps row predicates inside the loop (no SGPR masks)
      const int j = t - 64 * w - lane;
      const bool valid = j >= 0 && j <= Ly && i <= Lx;
      // ---- boundary row of the previous workgroup: fetch the next chunk
      if (xin && t <= Ly && (t % CQ) == 0) {
        const int e_end = min(t + CQ, Ly + 1);
        if (ok_run) ok_run = wait_ge(Pr.flagv + pd.fin, e_end, Pr.abortf, Pr.spin_max);
        fetch_chunk(s_xc, Pr.xring + pd.xin_off + (size_t)t * NX, (e_end - t) * NX);
      }
      // ---- emissions at (i, j): issued first, so that their LDS round trip
      // overlaps the ring read's
      int yj = dpp_i32<DPP_WAVE_SHR1>(yj_prev);
      if (lane == 0) yj = y0_next;
      double emM[NG], emI[NG];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        emM[g] = s_tab[g * TAB_SZ + xi * A_PAD + yj];
        emI[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + yj];
      }
      // ---- receive the predecessor row's h (lane-1, or a ring for lane 0)
      double rD[nD > 0 ? nD : 1], rM[nM > 0 ? nM : 1];
      int eR;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (L::type(b) == ST_D) rD[rank_in_type<L, ST_D>(b)] = dpp_f64<DPP_WAVE_SHR1>(hp[b]);
        if (L::type(b) == ST_M) rM[rank_in_type<L, ST_M>(b)] = dpp_f64<DPP_WAVE_SHR1>(hp[b]);
      }
      eR = dpp_i32<DPP_WAVE_SHR1>(eH);
      if (lane == 0 && valid && i >= 1) {
        const double* e = (w >= 1) ? s_ring + ((size_t)(w - 1) * 2 + (j & 1)) * NX
                                   : s_xc + (size_t)(j % CQ) * NX;
#pragma unroll
        for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
        for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
        eR = (int)e[NX - 1];
      }
      // ---- new cell: raw products and per-type maxima.  No validity masks:
      // a valid cell's out-of-lattice predecessors (row -1: DPP zero fill;
      // column -1: cells of negative j, computed from zeros) hold exactly 0,
      // and invalid cells past Lx or Ly feed only invalid cells.
      double f[NB];
      double mxM = 0.0, mxI = 0.0, mxD = 0.0;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b), g = L::group(b);
        double v;
        if (ty == ST_M) { v = rM_last[rank_in_type<L, ST_M>(b)] * emM[g]; mxM = fmax(mxM, v); }
        else if (ty == ST_I) { v = hp[b] * emI[g]; mxI = fmax(mxI, v); }
        else { v = rD[rank_in_type<L, ST_D>(b)] * emD[g]; mxD = fmax(mxD, v); }
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
      const bool origin = valid && i == 0 && j == 0;  // F = 0, E = 0 there already
      // ---- store the forward lattice (band-local step u = j + lane)
      if (want_lat && valid) {
        store_slot<NB, true>(lat + (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + 2 * lane, F, E);
      }
      // ---- push: h[k] = sum_s F[s] P[s][k]
      double h[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) h[k] = 0.0;
      row_contract<NB>(h, prow, F);
      if (pd.band0 == 0 && w == 0 && t == 0) {  // the origin's step (wave-uniform test)
        if (lane == 0) {
#pragma unroll
          for (int k = 0; k < NB; ++k) h[k] = p_start[k];
        }
      }
      // ---- corner: log P, its mantissa / exponent for the backward, the E column
      if (valid && i == Lx && j == Ly) {
        double tot = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b) tot = fma(F[b], p_end[b], tot);
        double lp;
        if (origin) { tot = p_end[NB]; lp = log(tot); }
        else lp = log(tot) + (double)E * LN2;
        const int ep = (tot > 0.0) ? E + frexp_exp(tot) : 0;
        const double mp = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
        Pr.out_logP[pd.pair] = (tot > 0.0) ? lp : -1e30;
        Pr.pinfo[2 * pd.pair] = mp;
        Pr.pinfo[2 * pd.pair + 1] = (double)ep;
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
        if (xout) {
          put_entry<L, NX>(Pr.xring + pd.xout_off + (size_t)j * NX, true, h, E);
          if (((j + 1) % CQ) == 0 || j == Ly) publish(Pr.flagv + pd.fout, j + 1);
        } else if (w < NW - 1) {
          put_entry<L, NX>(s_ring + ((size_t)w * 2 + (j & 1)) * NX, false, h, E);
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) hp[b] = h[b];
#pragma unroll
      for (int m = 0; m < nM; ++m) rM_last[m] = rM[m];
      eH = E;
      eM_last = eR;
      yj_prev = yj;
      y0_next = s_y[min(max(t - 64 * w, 0), Ly)];  // lane 0's y_{j-1} of the next step
    }
    lds_barrier();
  }
  if (!ok_run) {
    __hip_atomic_store(Pr.abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) Pr.out_logP[pd.pair] = __builtin_nan("");
  }
}

// ============================================================= backward
// W recursion waves (one band each) plus one count wave (w == W).
template <class L, int W>
__global__ __launch_bounds__((W + 1) * 64, 2) void fbc_bwd_kernel(CParams Pr, const double* __restrict__ pmat,
                                                               const CDesc* __restrict__ descs) {
  constexpr int NTHR = (W + 1) * 64;
  constexpr int NB = L::NB, NG = L::NG;
  constexpr int nM = count_type<L, ST_M>();
  constexpr int nD = count_type<L, ST_D>();
  using LY = CLds<L, W>;
  constexpr int NX = LY::NX;
  extern __shared__ double smem[];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const CDesc pd = descs[take_ticket(Pr.ticket, Pr.abortf)];
  const int Lx = pd.Lx, Ly = pd.Ly, NW = pd.nw;
  const int S = Pr.S;
  const int W1 = Ly + 1;
  const int BSTRIDE = Ly + 64;

  double* s_tab = smem + LY::tab;
  double* s_ecnt = smem + LY::ecnt;
  double* s_ring = smem + LY::ring(true);
  double* s_xc = smem + LY::xc(true);
  double* s_stage = smem + LY::stage(true);
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + LY::seq_bytes(true);
  uint8_t* s_y = s_x + LY::xbytes;
  __shared__ int s_cflag[W];  // count wave -> recursion wave w: stage read up to step s_cflag[w] - 1
  stage_inputs<L, W, NTHR>(Pr, pd, pmat + pd.pm_off, s_tab, s_x, s_y);
  if (threadIdx.x < W) s_cflag[threadIdx.x] = 0;
  const bool want_ecnt = (Pr.flags & TKFMI_WANT_EMIT_COUNTS) && Pr.out_ecnt != nullptr;
  if (want_ecnt)
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) s_ecnt[q] = 0.0;

  const double* __restrict__ pm = pmat + pd.pm_off;
  const double* __restrict__ Pf = pm;
  const double* p_start = smem + LY::misc;
  const double* p_end = p_start + NB;
  double prowT[NB];  // prowT[k'] in lane l: P[l % 16][k'] (0 past NB)
#pragma unroll
  for (int k2 = 0; k2 < NB; ++k2)
    prowT[k2] = valu_copy((lane & 15) < NB ? pm[NB * NB + k2 * NB + (lane & 15)] : 0.0);

  const bool want_counts = !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  const bool want_post = (Pr.flags & TKFMI_WANT_POSTERIORS) && Pr.out_post != nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  const double* lat = Pr.lattice + pd.lat_off;
  const double mP = Pr.pinfo[2 * pd.pair];
  const int EP = (int)Pr.pinfo[2 * pd.pair + 1];
  if (!(mP > 0.0)) return;  // impossible pair (uniform over the workgroup and its chain)
  const double inv_mP = 1.0 / mP;
  const bool has_band = w < NW;
  const int band = pd.band0 + NW - 1 - w;  // wave 0 holds the bottom band
  const int i_band = 64 * band + lane;
  const bool xin = (w == 0) && pd.fin >= 0;
  const bool xout = (w == NW - 1) && pd.fout >= 0;
  bool ok_run = true;
  const bool counter = w == W;
  d4 cacc = {0.0, 0.0, 0.0, 0.0}, cacc2 = {0.0, 0.0, 0.0, 0.0};  // count wave's accumulators
  // Count staging, one step deep: at the end of step t a recursion wave
  // stages its cells in its buffer (after the count wave has read the cells of
  // step t-1 from it); during step t+1 the count wave runs their MFMAs.
  double* stage_f = s_stage + (size_t)(counter ? 0 : w) * kStageDoubles;  // [64 cells][16] F-scaled
  double* stage_g = stage_f + kStageDoubles / 2;                          // [64 cells][16] g
  double accD[NG];   // D-type emission mass of this lane's row (x fixed per lane)
  double runI[NG];   // I-type mass of the column passing up through the lanes
#pragma unroll
  for (int gq = 0; gq < NG; ++gq) { accD[gq] = 0.0; runI[gq] = 0.0; }
  double* post = want_post ? Pr.out_post + pd.post_off : nullptr;
  __syncthreads();

  double qp[NB];
  double rM_last[nM];
#pragma unroll
  for (int b = 0; b < NB; ++b) qp[b] = 0.0;
#pragma unroll
  for (int m = 0; m < nM; ++m) rM_last[m] = 0.0;
  int eQ = EXP_NONE, eM_last = EXP_NONE;
  // x fixed per lane; y passes up one lane per step (lane 63 reads ahead).
  const int xi = s_x[64 * max(NW - 1 - w, 0) + lane];
  double emD[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) emD[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + A_PAD + xi];
  int yj_prev = 0;
  int y63_next = s_y[max(Ly - 1, 0)];
  const int off = 64 * w + (63 - lane);  // position p = t - off, column j = Ly - p
  // Lattice slot of this lane's cell at position p.  An invalid cell reads
  // the origin's slot (cell (0, 0), always written by the forward, so finite);
  // its count operands are then scaled by 0.
  auto f_addr = [&](int p, int i) -> const double* {
    const bool v = has_band && p >= 0 && p <= Ly && i <= Lx;
    return lat + (v ? (((size_t)band * BSTRIDE + (Ly - p + lane)) * SLOTS) * 64 + 2 * lane : 0);
  };
  const int T = pd.T;
  // A wave's band is in range for steps 64 w .. 64 w + 63 + Ly; outside
  // them every lane is invalid and the wave only keeps the barrier (its
  // carried values are zero there, as the skipped steps would leave them).
  const int t_on = 64 * w, t_off = 64 * w + 63 + Ly;
  // The corner (Lx, Ly) is in wave 0 of the chain's last workgroup, lane
  // Lx - 64 band, at position 0.
  const bool corner_wave = has_band && band == (Lx + 64) / 64 - 1;
  const int t_corner = 64 * w + 63 - (Lx - 64 * band);
  for (int t = 0; t < T; ++t) {
    if (counter && want_counts) count_step<W>(s_stage, NW, Ly, t, lane, s_cflag, cacc, cacc2);
    if (has_band && t >= t_on && t <= t_off) {
      int i = i_band;
      asm volatile("" : "+v"(i));  // keeps row predicates inside the loop (no SGPR masks)
      const int p = t - off;
      const int j = Ly - p;
      const bool valid = p >= 0 && p <= Ly && i <= Lx;
      // ---- boundary row of the next workgroup (below): fetch the next chunk
      // (before this step's lattice loads: loads complete in order, so the
      // chunk's wait would otherwise include them)
      if (xin && t <= Ly && (t % CQ) == 0) {
        const int e_end = min(t + CQ, Ly + 1);
        if (ok_run) ok_run = wait_ge(Pr.flagv + pd.fin, e_end, Pr.abortf, Pr.spin_max);
        fetch_chunk(s_xc, Pr.xring + pd.xin_off + (size_t)t * NX, (e_end - t) * NX);
      }
      // The cell's forward values: loaded here, first used after the
      // transition contraction, which covers the load latency.
      double fcur[SLOTS];
      {
        const double* a = f_addr(p, i);
        load_slot<NB, true>(a, fcur);
      }
      // ---- receive the successor row's q (lane+1, or a ring for lane 63)
      double rD[nD > 0 ? nD : 1], rM[nM > 0 ? nM : 1];
      int eR;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (L::type(b) == ST_D) rD[rank_in_type<L, ST_D>(b)] = dpp_f64<DPP_WAVE_SHL1>(qp[b]);
        if (L::type(b) == ST_M) rM[rank_in_type<L, ST_M>(b)] = dpp_f64<DPP_WAVE_SHL1>(qp[b]);
      }
      eR = dpp_i32<DPP_WAVE_SHL1>(eQ);
      if (lane == 63 && valid && i < Lx) {
        const double* e = (w >= 1) ? s_ring + ((size_t)(w - 1) * 2 + (p & 1)) * NX
                                   : s_xc + (size_t)(p % CQ) * NX;
#pragma unroll
        for (int m = 0; m < nM; ++m) rM[m] = e[m];
#pragma unroll
        for (int d = 0; d < nD; ++d) rD[d] = e[nM + d];
        eR = (int)e[NX - 1];
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
      row_contract<NB>(Bv, prowT, g);
      const bool corner = valid && i == Lx && j == Ly;
      int EBraw = Eg;
      if (corner_wave && t == t_corner) {  // the corner's step (wave-uniform test)
        if (corner) {
#pragma unroll
          for (int k = 0; k < NB; ++k) Bv[k] = p_end[k];
          EBraw = 0;
        }
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
      if (want_counts) {  // stage this step's cells (counted during the next step)
        int spins = 0;
        while (__hip_atomic_load(s_cflag + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < t) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > kSpinMax) { ok_run = false; break; }
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) {
          const int o = lane * 16 + (s ^ (lane & 15));
          stage_f[o] = (s < NB) ? fcur[s < NB ? s : 0] * fscale : 0.0;
          stage_g[o] = (s < NB) ? g[s < NB ? s : 0] : 0.0;
        }
      }
      if (want_counts) {
        if (valid && i == 0 && j == 0) {  // start state row: n[S, k]
          const double sc = ldexp(inv_mP, Eg - EP);
#pragma unroll
          for (int b = 0; b < NB; ++b) out_n[b + L::SOFF] = p_start[b] * g[b] * sc;
        }
      }
      int yj = dpp_i32<DPP_WAVE_SHL1>(yj_prev);
      if (lane == 63) yj = y63_next;
      double emM[NG], emI[NG];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        emM[g] = s_tab[g * TAB_SZ + xi * A_PAD + yj];
        emI[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + yj];
      }
      // ---- expected emission counts: posterior mass per (group, type)
      // M-type mass goes to its (x, y) entry with an LDS atomic; D-type mass
      // accumulates per lane (the row, hence x, is fixed); I-type mass rides
      // a running sum up the column (lane l+1 -> l, as q does) and lane 0
      // adds the column total of the band to its y entry.
      if (want_ecnt) {
        double cI[NG];
#pragma unroll
        for (int gq = 0; gq < NG; ++gq) { runI[gq] = dpp_f64<DPP_WAVE_SHL1>(runI[gq]); cI[gq] = 0.0; }
        if (valid) {
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
            if (i >= 1 && j >= 1) atomicAdd(s_ecnt + gq * TAB_SZ + xi * A_PAD + yj, sM[gq] * pscale);
            if (j >= 1) cI[gq] = sI[gq] * pscale;
            if (i >= 1) accD[gq] = fma(sD[gq], pscale, accD[gq]);
          }
        }
#pragma unroll
        for (int gq = 0; gq < NG; ++gq) runI[gq] += cI[gq];
        if (lane == 0 && valid && j >= 1) {
#pragma unroll
          for (int gq = 0; gq < NG; ++gq) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + yj, runI[gq]);
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
      // No masks: an invalid cell's g is 0 (its ok flags are false), so its q
      // is 0; a valid cell in row 0 or column 0 gets a q for the types whose
      // emission is not allowed there, but that q feeds only cells outside
      // the lattice (row -1, column -1), which no valid cell reads.
      double q[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b), gq = L::group(b);
        const double em = ty == ST_M ? emM[gq] : (ty == ST_I ? emI[gq] : emD[gq]);
        q[b] = em * Bv[b];
      }
      // ---- publish this row's q for the band above (lane 0)
      if (lane == 0 && valid) {
        if (xout) {
          put_entry<L, NX>(Pr.xring + pd.xout_off + (size_t)p * NX, true, q, EB);
          if (((p + 1) % CQ) == 0 || p == Ly) publish(Pr.flagv + pd.fout, p + 1);
        } else if (w < NW - 1) {
          put_entry<L, NX>(s_ring + ((size_t)w * 2 + (p & 1)) * NX, false, q, EB);
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) qp[b] = q[b];
#pragma unroll
      for (int m = 0; m < nM; ++m) rM_last[m] = rM[m];
      eQ = EB;
      eM_last = eR;
      yj_prev = yj;
      y63_next = s_y[min(max(Ly - (t + 1 - 64 * w) - 1, 0), Ly)];  // lane 63's y_{j-1} of the next step
    }
    lds_barrier();
  }
  if (counter && want_counts) count_step<W>(s_stage, NW, Ly, T, lane, s_cflag, cacc, cacc2);  // last step's cells
  if (has_band && want_ecnt && i_band >= 1 && i_band <= Lx) {
#pragma unroll
    for (int gq = 0; gq < NG; ++gq)
      if (accD[gq] != 0.0) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + A_PAD + xi, accD[gq]);
  }
  __syncthreads();
  if (!ok_run) {
    __hip_atomic_store(Pr.abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) Pr.out_logP[pd.pair] = __builtin_nan("");
  }
  if (want_ecnt) {  // layout group g -> caller group gmap[g]
    double* oe = Pr.out_ecnt + (size_t)pd.pair * Pr.G_tab * TAB_SZ;
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) {
      const int gq = q / TAB_SZ, o = q - gq * TAB_SZ;
      const double v = s_ecnt[q];
      if (v != 0.0) unsafeAtomicAdd(oe + Pr.gmap[gq] * TAB_SZ + o, v);
    }
  }
  if (!want_counts) return;
  // ---- the count wave's accumulators through LDS, then add
  double* s_red = s_stage;  // [16][16] (the count wave has read every stage)
  if (counter) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * v + (lane >> 4), col = lane & 15;  // accumulator layout
      s_red[row * 16 + col] = cacc[v] + cacc2[v];
    }
  }
  __syncthreads();
  for (int qq = threadIdx.x; qq < NB * NB; qq += NTHR) {
    const int k = qq / NB, k2 = qq % NB;
    const double acc = s_red[k * 16 + k2];
    const double pr = Pf[k * NB + k2];
    if (pr > 0.0 && acc != 0.0) unsafeAtomicAdd(out_n + (k + L::SOFF) * S + (k2 + L::SOFF), pr * acc);
  }
}

__global__ void zero_ints_kernel(int32_t* p, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) p[q] = 0;
}

// ---------------------------------------------------------------- host side
// Launch classes, each with its own stream: 0 = the longest pairs, 1 = the
// other pairs of at least main_w bands, 2 = 2-wave, 3 = 1-wave workgroups.
// The longest pairs set the makespan (their backward waits for their class's
// forward launch), so they get a class of their own that starts first.
// Waves per workgroup of classes 0 and 1 are tuning parameters
// (tkfmi_chain_config; tkfmi.autotune picks them per workload): a 4-wave
// backward workgroup needs ~92 KB of LDS (one per compute unit), a 3-wave
// one ~76 KB (two fit).
constexpr int NCLS = 4;
struct ChainConfig {
  int w[NCLS] = {3, 3, 2, 1};
  // cells >= long_frac x the batch's largest pair; 0 = no long class
  double long_frac = 0.7;
  // 1 (default): the long class runs on the one-launch schedule
  // (fb2d_bidi.hip) while the other classes keep two launches, so the
  // longest pairs stop setting the call's length (the 1024-pair K=3 batch:
  // 8.1-8.3 ms with its 6-15 longest pairs so, against 8.8 ms with no long
  // class; a long class in two launches does not shorten the call)
  int long_one = 1;
};
std::mutex g_cfg_mu;
ChainConfig g_cfg = [] {
  ChainConfig c;
  if (const char* e = getenv("TKFMI_CHAIN_LONG")) c.long_frac = atof(e);
  if (const char* e = getenv("TKFMI_CHAIN_LONG_ONE")) c.long_one = atoi(e) != 0;
  return c;
}();
// The one-launch schedule (fb2d_bidi.hip): a pair's forward and backward
// meet at C = cut (Lx + Ly); the forward counts the cells past C, the
// backward those up to it.  It halves a pair's critical path but keeps all of
// a pair's workgroups resident at once, and the workgroups of a chain's
// later links then wait for their producers in slots that other pairs could
// use: on a GPU it saturates it is slower than two launches (the 1024-pair
// K=3 batch: 12.1 against 8.9 ms), on one it leaves room it is faster (the
// batch's longest pair alone: 4.0 against 6.8 ms).  So it runs when every
// workgroup of the call fits the GPU at once: at most max_wg workgroups (2
// per band group of each pair; -1: 2 per compute unit).  cut = 0: never.
// tkfmi_chain_schedule, TKFMI_CHAIN_CUT, TKFMI_CHAIN_BIDI_MAX_WG.
std::atomic<double> g_cut{[] {
  const char* e = getenv("TKFMI_CHAIN_CUT");
  return e ? std::max(0.0, std::min(1.0, atof(e))) : 0.45;
}()};
std::atomic<int> g_bidi_max_wg{[] {
  const char* e = getenv("TKFMI_CHAIN_BIDI_MAX_WG");
  return e ? atoi(e) : -1;
}()};
// Workgroups of the one-launch kernel that the current device holds at once
// (two per compute unit: VGPR- and LDS-bound).
int bidi_resident_wg() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 0;
  return 2 * cus;
}
// Configurations chosen by tkfmi_autotune, per (layout, length bin of the
// batch's longest sequence); a batch with no entry uses g_cfg.
constexpr int kLenBins = 6;  // longest sequence < 128, 256, 512, 1024, 2048, more
std::map<std::pair<int, int>, ChainConfig> g_tuned;
int length_bin(int max_len) {
  int b = 0;
  for (int lim = 128; b < kLenBins - 1 && max_len >= lim; lim *= 2) ++b;
  return b;
}
int max_seq_len(const tkfmi_problem* p) {
  int m = 0;
  for (int b = 0; b < p->B; ++b) m = std::max(m, std::max(p->h_x_len[b], p->h_y_len[b]));
  return m;
}
ChainConfig chain_config(int layout, int bin) {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  auto it = g_tuned.find({layout, bin});
  return it != g_tuned.end() ? it->second : g_cfg;
}
// Class of a pair with nb bands and `cells` lattice cells.
int ccls_of(const ChainConfig& cf, int nb, int64_t cells, int64_t cells_max) {
  if (nb >= cf.w[0] && cf.w[0] > 1 && cf.long_frac > 0.0 && (double)cells >= cf.long_frac * (double)cells_max)
    return 0;
  if (nb >= cf.w[1] && cf.w[1] >= 2) return 1;
  if (nb >= 2) return 2;
  return 3;
}

// Ticket order of each class: 1 (default) = critical path first, 0 = pair
// by pair (largest pairs first, a pair's chain consecutive).
// TKFMI_CHAIN_ORDER.
int chain_order() {
  static const int v = [] {
    const char* e = getenv("TKFMI_CHAIN_ORDER");
    return e ? atoi(e) : 1;
  }();
  return v;
}

struct ChainPlan {
  bool ok = false;
  std::string why;
  std::vector<CDesc> fwd[NCLS], bwd[NCLS];
  std::vector<CDesc> both[NCLS];  // one-launch schedule: each pair's forward and backward groups together
  bool bidi = false;              // some class runs on the one-launch schedule
  bool bidi_cls[NCLS] = {false, false, false, false};
  int fail_base = 0;  // one-launch schedule: flag index of pair 0's failure flag
  int64_t cls_cells[NCLS] = {0, 0, 0, 0};
  int cls_pairs[NCLS] = {0, 0, 0, 0};
  int max_ly[NCLS] = {0, 0, 0, 0};
  size_t desc_b = 0, pm_b = 0, lat_b = 0, ring_b = 0, flag_b = 0, pinfo_b = 0, cut_b = 0, pslot_b = 0;
  // ints: NCLS classes x 2 tickets, abort, pad, then 2 per boundary (and,
  // one-launch, per pair 1 partial counter + 1 cut flag per band group)
  int n_flags = 0;
  int W[NCLS] = {3, 3, 2, 1};  // waves per workgroup of each class
  size_t total() const { return desc_b + pm_b + lat_b + ring_b + flag_b + pinfo_b + cut_b + pslot_b; }
};

// The one-launch schedule applies when the backward has counts to take, no
// posterior lattice is asked for and the call's n_wg workgroups fit the
// limit (see g_cut).
bool bidi_wanted(uint32_t flags, int64_t n_wg) {
  if (flags & (TKFMI_FORWARD_ONLY | TKFMI_WANT_POSTERIORS)) return false;
  if ((flags & TKFMI_NO_COUNTS) && !(flags & TKFMI_WANT_EMIT_COUNTS)) return false;
  if (!(g_cut.load() > 0.0)) return false;
  const int lim = g_bidi_max_wg.load();
  return n_wg <= (int64_t)(lim < 0 ? bidi_resident_wg() : lim);
}

// cfg: the class configuration to plan with (null: the tuned or default one
// for the batch's layout and length bin).
ChainPlan plan_chain(const tkfmi_problem* p, const Plan& pl, uint32_t flags, const ChainConfig* cfg = nullptr) {
  ChainPlan cp;
  const int B = p->B;
  const int NX = nx_of(pl);
  std::vector<int> order(B);
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0 || lx + ly > 32000) { cp.why = "sequence lengths out of range"; return cp; }
    order[b] = b;
  }
  // longest chains first inside each class (the ticket order)
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int64_t ca = (int64_t)(p->h_x_len[a] + 1) * (p->h_y_len[a] + 1);
    const int64_t cb = (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
    return ca > cb;
  });
  std::vector<int64_t> lat_off(B), post_off(B);
  int64_t al = 0, ap = 0;
  for (int b = 0; b < B; ++b) {
    lat_off[b] = al;
    if (!(flags & TKFMI_FORWARD_ONLY)) al += (int64_t)((p->h_x_len[b] + 64) / 64) * (p->h_y_len[b] + 64) * SLOTS * 64;
    post_off[b] = ap;
    ap += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1) * p->S;
  }
  int64_t aring = 0, acut = 0, aslot = 0;
  int nfl = 2 * NCLS + 8;
  const double cut = g_cut.load();
  std::vector<std::pair<int64_t, std::vector<CDesc>>> pair_groups[NCLS];  // one-launch: (critical path, groups)
  const int64_t cells_max = B > 0 ? (int64_t)(p->h_x_len[order[0]] + 1) * (p->h_y_len[order[0]] + 1) : 0;
  ChainConfig cf = cfg ? *cfg : chain_config((int)pl.lay, length_bin(max_seq_len(p)));
  for (int c = 0; c < NCLS; ++c) cp.W[c] = cf.w[c];
  {
    int64_t n_wg = 0;  // workgroups of a one-launch call: 2 per band group
    for (int b = 0; b < B; ++b) {
      const int nb = (p->h_x_len[b] + 64) / 64;
      const int c = ccls_of(cf, nb, (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1), cells_max);
      n_wg += 2 * ((nb + cf.w[c] - 1) / cf.w[c]);
    }
    const bool all = bidi_wanted(flags, n_wg);
    // A long class on one launch exists only in calls too large for the
    // whole call to run on it: there the longest pairs' chains set the
    // call's length (c1-shaped batches of 200-1024 pairs, 630-3300
    // workgroups: 4.1-8.2 ms against 5.1-8.4 with every pair on two
    // launches and 4.6-12.2 with every pair on one); in calls that fit, its
    // pairs stay in the main class (a class of its own, on a stream of its
    // own, made the serving loop's latency worse).
    const bool long_one = cf.long_one && cf.long_frac > 0.0 && !all && bidi_wanted(flags, 0);
    if (cf.long_one && !long_one) cf.long_frac = 0.0;
    for (int c = 0; c < NCLS; ++c) cp.bidi_cls[c] = all || (c == 0 && long_one);
    cp.bidi = cp.bidi_cls[0] || cp.bidi_cls[1] || cp.bidi_cls[2] || cp.bidi_cls[3];
  }
  for (int q = 0; q < B; ++q) {
    const int b = order[q];
    const int Lx = p->h_x_len[b], Ly = p->h_y_len[b];
    const int nb = (Lx + 64) / 64;
    const int c = ccls_of(cf, nb, (int64_t)(Lx + 1) * (Ly + 1), cells_max);
    const int W = cf.w[c];
    const int G = (nb + W - 1) / W;
    ++cp.cls_pairs[c];
    cp.cls_cells[c] += (int64_t)(Lx + 1) * (Ly + 1);
    cp.max_ly[c] = std::max(cp.max_ly[c], Ly);
    const int64_t ring0 = aring;
    const int flag0 = nfl;
    // the backward's boundary rows: the same region as the forward's when
    // the two run in turn, a region of their own when they run together
    const bool bidi = cp.bidi_cls[c];
    const int64_t bring0 = bidi ? ring0 + (int64_t)(G - 1) * (Ly + 1) * NX : ring0;
    aring += (int64_t)(bidi ? 2 : 1) * (G - 1) * (Ly + 1) * NX;
    nfl += 2 * (G - 1);
    std::vector<CDesc> ch(G);
    for (int g = 0; g < G; ++g) {
      CDesc d;
      memset(&d, 0, sizeof(d));
      d.pair = b; d.Lx = Lx; d.Ly = Ly; d.band0 = g * W;
      d.nw = std::min(W, nb - g * W); d.g = g; d.G = G;
      d.x_off = p->h_x_off[b]; d.y_off = p->h_y_off[b];
      d.lat_off = lat_off[b]; d.post_off = post_off[b];
      d.pm_off = (p->n_trans > 1 ? (int64_t)b : 0) * (int64_t)stride_for(pl.NB);
      ch[g] = d;
    }
    for (int g = 0; g < G; ++g) {  // forward: boundary g-1 in, g out
      CDesc d = ch[g];
      const int rl = std::min(63, Lx - 64 * (d.band0 + d.nw - 1));
      d.T = Ly + 64 * (d.nw - 1) + rl + 1;
      d.xin_off = g > 0 ? ring0 + (int64_t)(g - 1) * (Ly + 1) * NX : -1;
      d.fin = g > 0 ? flag0 + 2 * (g - 1) : -1;
      d.xout_off = g < G - 1 ? ring0 + (int64_t)g * (Ly + 1) * NX : -1;
      d.fout = g < G - 1 ? flag0 + 2 * g : -1;
      cp.fwd[c].push_back(d);
    }
    for (int g = G - 1; g >= 0; --g) {  // backward: boundary g in, g-1 out
      CDesc d = ch[g];
      d.T = Ly + 64 * (d.nw - 1) + 64;
      d.xin_off = g < G - 1 ? bring0 + (int64_t)g * (Ly + 1) * NX : -1;
      d.fin = g < G - 1 ? flag0 + 2 * g + 1 : -1;
      d.xout_off = g > 0 ? bring0 + (int64_t)(g - 1) * (Ly + 1) * NX : -1;
      d.fout = g > 0 ? flag0 + 2 * (g - 1) + 1 : -1;
      cp.bwd[c].push_back(d);
    }
    if (bidi) {
      // the meeting antidiagonal, the pair's partial counter, one cut flag
      // per band group, partial slots (16 doubles apart: one cache line
      // each) and the cut buffer ([2][16][Lx + 1] doubles)
      const int D = Lx + Ly;
      const int Cm = std::max(0, std::min(D, (int)llround(cut * D)));
      const int cnt = nfl++;
      const int bf0 = nfl;
      nfl += G;
      const int ps0 = (int)aslot;
      aslot += 16 * G;
      const int64_t cut0 = acut;
      acut += 32 * (int64_t)(Lx + 1);
      std::vector<CDesc>& fw = cp.fwd[c];
      std::vector<CDesc>& bw = cp.bwd[c];
      std::vector<CDesc> grp;
      for (int k = 0; k < G; ++k) {  // chain heads first: forward 0, backward G-1, forward 1, ...
        for (int dir = 0; dir < 2; ++dir) {
          CDesc d = dir ? bw[bw.size() - G + k] : fw[fw.size() - G + k];
          d.dir = dir;
          d.C = Cm;
          d.cnt = cnt;
          d.bflag = bf0 + d.g;
          d.pslot = ps0;
          d.gpair = G;
          d.cut_off = cut0;
          grp.push_back(d);
        }
      }
      const int lag = 64 * (cp.W[c] - 1) + 64 + CQ;
      const CDesc& f0 = fw[fw.size() - G];
      pair_groups[c].push_back({(int64_t)(G - 1) * lag + f0.T, grp});
      fw.resize(fw.size() - G);
      bw.resize(bw.size() - G);
    }
  }
  if (cp.bidi) {  // per pair, a flag raised by its timed-out waits
    cp.fail_base = nfl;
    nfl += B;
  }
  if (cp.bidi) {  // pairs by critical path, each pair's groups consecutive
    for (int c = 0; c < NCLS; ++c) {
      std::stable_sort(pair_groups[c].begin(), pair_groups[c].end(),
                       [](const auto& a, const auto& b) { return a.first > b.first; });
      for (auto& pg : pair_groups[c]) cp.both[c].insert(cp.both[c].end(), pg.second.begin(), pg.second.end());
    }
  }
  // Critical-path ticket order.  A workgroup's critical path is the number
  // of steps from its start to the end of its pair's chain, (links after it)
  // x lag + T, where lag is how far a consumer's first step trails its
  // producer's (the producer's last wave is 64 (W - 1) steps behind its
  // first, plus one chunk).  Longest first: a long pair's later links are
  // dispatched early enough to keep up with their producers, and the short
  // pairs fill in behind.  The critical path falls along a chain, so
  // producers keep lower tickets than their consumers.
  if (chain_order() == 1) {
    for (int c = 0; c < NCLS; ++c) {
      const int lag = 64 * (cp.W[c] - 1) + 64 + CQ;
      for (int dir = 0; dir < 2; ++dir) {
        std::vector<CDesc>& v = dir ? cp.bwd[c] : cp.fwd[c];
        auto crit = [&](const CDesc& d) {
          const int links = dir ? d.g : d.G - 1 - d.g;
          return (int64_t)links * lag + d.T;
        };
        std::stable_sort(v.begin(), v.end(), [&](const CDesc& a, const CDesc& b) { return crit(a) > crit(b); });
      }
    }
  }
  size_t nd = 0;
  for (int c = 0; c < NCLS; ++c) nd += cp.fwd[c].size() + cp.bwd[c].size() + cp.both[c].size();
  cp.desc_b = align256(sizeof(CDesc) * std::max<size_t>(nd, 1));
  cp.pm_b = align256(stride_for(pl.NB) * 8 * (size_t)p->n_trans);
  cp.lat_b = align256((size_t)al * 8);
  cp.ring_b = align256((size_t)std::max<int64_t>(aring, 1) * 8);
  cp.n_flags = nfl;
  cp.flag_b = align256((size_t)nfl * 4);
  cp.pinfo_b = align256((size_t)std::max(B, 1) * 16);
  cp.cut_b = cp.bidi ? align256((size_t)std::max<int64_t>(acut, 1) * 8) : 0;
  cp.pslot_b = cp.bidi ? align256((size_t)std::max<int64_t>(aslot, 1) * 8) : 0;
  for (int c = 0; c < NCLS; ++c)
    if (cp.cls_pairs[c] && CLds<LMix<3, 1>, 4>::total(true, cp.max_ly[c]) > 160 * 1024 - 512) {
      cp.why = "LDS budget exceeded (sequences too long)";
      return cp;
    }
  cp.ok = true;
  return cp;
}

// Workgroups per CU that each class's launches leave room for (through their
// LDS request); 0 = as many as fit.  TKFMI_CHAIN_WPC="f0,b0,f1,b1,f2,b2,f3,b3".
void chain_wpc(int wpc[NCLS][2]) {
  static int v[2 * NCLS] = {0, 0, 0, 0, 0, 0, 0, 0};
  static bool init = false;
  if (!init) {
    const char* e = getenv("TKFMI_CHAIN_WPC");
    if (e) sscanf(e, "%d,%d,%d,%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
    init = true;
  }
  for (int c = 0; c < NCLS; ++c) { wpc[c][0] = v[2 * c]; wpc[c][1] = v[2 * c + 1]; }
}

template <class L, int W>
int launch_chain(bool bwd, const CParams& prm, const double* pmat, const CDesc* d, int n, int max_ly,
                 int wpc, hipStream_t s, std::string& err) {
  const void* fn = bwd ? (const void*)fbc_bwd_kernel<L, W> : (const void*)fbc_fwd_kernel<L, W>;
  size_t smem = CLds<L, W>::total(bwd, max_ly);
  if (wpc > 0) smem = std::max(smem, (size_t)(160 * 1024 / wpc - 512) & ~(size_t)255);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  if (bwd) hipLaunchKernelGGL((fbc_bwd_kernel<L, W>), dim3(n), dim3((W + 1) * 64), smem, s, prm, pmat, d);
  else hipLaunchKernelGGL((fbc_fwd_kernel<L, W>), dim3(n), dim3(W * 64), smem, s, prm, pmat, d);
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("fbc launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <class L>
int launch_chain_cls(int W, bool bwd, const CParams& prm, const double* pmat, const CDesc* d, int n,
                     int max_ly, int wpc, hipStream_t s, std::string& err) {
  switch (W) {
    case 4: return launch_chain<L, 4>(bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case 3: return launch_chain<L, 3>(bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case 2: return launch_chain<L, 2>(bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    default: return launch_chain<L, 1>(bwd, prm, pmat, d, n, max_ly, wpc, s, err);
  }
}

int dispatch_chain(const Plan& pl, int W, bool bwd, const CParams& prm, const double* pmat,
                   const CDesc* d, int n, int max_ly, int wpc, hipStream_t s, std::string& err) {
  switch (pl.lay) {
    case Layout::Tkf: return launch_chain_cls<LTkf>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix11: return launch_chain_cls<LMix<1, 1>>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix21: return launch_chain_cls<LMix<2, 1>>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix31: return launch_chain_cls<LMix<3, 1>>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix12: return launch_chain_cls<LMix<1, 2>>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix13: return launch_chain_cls<LMix<1, 3>>(W, bwd, prm, pmat, d, n, max_ly, wpc, s, err);
    default: err = "layout not handled"; return TKFMI_EINVAL;
  }
}

int prep_dispatch(const Plan& pl, const tkfmi_problem* p, double* pmat, hipStream_t s, std::string& err) {
  switch (pl.lay) {
    case Layout::Tkf: return launch_prep<LTkf>(p, pmat, s, err);
    case Layout::Mix11: return launch_prep<LMix<1, 1>>(p, pmat, s, err);
    case Layout::Mix21: return launch_prep<LMix<2, 1>>(p, pmat, s, err);
    case Layout::Mix31: return launch_prep<LMix<3, 1>>(p, pmat, s, err);
    case Layout::Mix12: return launch_prep<LMix<1, 2>>(p, pmat, s, err);
    case Layout::Mix13: return launch_prep<LMix<1, 3>>(p, pmat, s, err);
    default: err = "layout not handled"; return TKFMI_EINVAL;
  }
}

}  // namespace tkfmi_v2

using namespace tkfmi_v2;

bool tkfmi_chain_workspace(const tkfmi_problem* p, uint32_t flags, size_t* bytes) {
  Plan pl;
  if (!p || p->B < 0 || !p->state_types || !p->state_group || !detect(p, pl)) return false;
  if (p->B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off)) return false;
  if (p->n_trans != 1 && p->n_trans != p->B) return false;
  const ChainPlan cp = plan_chain(p, pl, flags);
  if (!cp.ok) return false;
  *bytes = cp.total();
  return true;
}

namespace {
int chain_run_impl(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err, const ChainConfig* cfg);
}

int tkfmi_chain_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                    hipStream_t stream, uint32_t flags, std::string& err) {
  return chain_run_impl(p, o, ws, ws_bytes, stream, flags, err, nullptr);
}

namespace {
int chain_run_impl(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err, const ChainConfig* cfg) {
  Plan pl;
  if (!detect(p, pl)) { err = "layout not handled"; return TKFMI_EINVAL; }
  const ChainPlan cp = plan_chain(p, pl, flags, cfg);
  if (!cp.ok) { err = cp.why; return TKFMI_EINVAL; }
  if (!ws || ws_bytes < cp.total()) {
    err = "workspace too small: need " + std::to_string(cp.total());
    return TKFMI_EWORKSPACE;
  }
  if (p->B == 0) return TKFMI_OK;
  char* wsb = static_cast<char*>(ws);
  CDesc* ddesc = reinterpret_cast<CDesc*>(wsb);
  double* pmat = reinterpret_cast<double*>(wsb + cp.desc_b);
  double* lattice = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b);
  double* xring = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b);
  int32_t* flagv = reinterpret_cast<int32_t*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b + cp.ring_b);
  double* pinfo = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b + cp.ring_b + cp.flag_b);
  double* cutbuf = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b + cp.ring_b + cp.flag_b + cp.pinfo_b);
  double* pslot = reinterpret_cast<double*>(wsb + cp.desc_b + cp.pm_b + cp.lat_b + cp.ring_b + cp.flag_b +
                                            cp.pinfo_b + cp.cut_b);
  // descriptors: [class][fwd | bwd] (two launches) or [class][both] (one launch)
  std::vector<CDesc> all;
  size_t base[NCLS][2];
  for (int c = 0; c < NCLS; ++c) {
    base[c][0] = all.size();
    all.insert(all.end(), cp.fwd[c].begin(), cp.fwd[c].end());
    all.insert(all.end(), cp.both[c].begin(), cp.both[c].end());
    base[c][1] = all.size();
    all.insert(all.end(), cp.bwd[c].begin(), cp.bwd[c].end());
  }
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
  rc = (flags & TKFMI_REUSE_DESCRIPTORS) ? TKFMI_OK : tkfmi_stage_upload(ddesc, all.data(), sizeof(CDesc) * all.size(), stream, err);
  if (rc != TKFMI_OK) return rc;
  hipLaunchKernelGGL(zero_ints_kernel, dim3(std::min(1024, (cp.n_flags + 255) / 256)), dim3(256), 0, stream,
                     flagv, cp.n_flags);
  if (hipGetLastError() != hipSuccess) { err = "zero launch failed"; return TKFMI_EHIP; }
  rc = prep_dispatch(pl, p, pmat, stream, err);
  if (rc != TKFMI_OK) return rc;
  if (!one_stream && hipEventRecord(sts->start, stream) != hipSuccess) { err = "hipEventRecord failed"; return TKFMI_EHIP; }
  CParams prm;
  memset(&prm, 0, sizeof(prm));
  prm.S = p->S; prm.flags = (int32_t)flags; prm.n_tab = p->n_tab; prm.G_tab = p->n_groups;
  for (int g = 0; g < 16; ++g) prm.gmap[g] = pl.gmap[g];
  prm.log_tab = p->log_tab; prm.x_cat = p->x_cat; prm.y_cat = p->y_cat;
  prm.lattice = lattice; prm.xring = xring; prm.flagv = flagv; prm.abortf = flagv + 2 * NCLS;
  prm.pinfo = pinfo;
  prm.cutbuf = cutbuf;
  prm.pslot = pslot;
  prm.fail_base = cp.fail_base;
  {
    const char* e = getenv("TKFMI_CHAIN_SPIN_LIMIT");
    prm.spin_max = e ? std::max(0, atoi(e)) : kSpinMax;
  }
  prm.out_logP = o->log_prob;
  prm.out_n = (flags & TKFMI_NO_COUNTS) ? nullptr : o->n_trans;
  prm.out_post = o->posteriors;
  prm.out_ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) ? o->emit_counts : nullptr;
  int wpc[NCLS][2];
  chain_wpc(wpc);
  const bool fwd_only = (flags & TKFMI_FORWARD_ONLY) != 0;
  for (int c = 0; c < NCLS; ++c) {
    if (!cp.cls_pairs[c]) continue;
    // Classes 2 and 3 (2- and 1-band pairs, short) share a library stream:
    // with the caller's stream that keeps a call within the GPU's 4 hardware
    // queues, so the long class's launch never shares one with the main
    // class's (two streams on one queue run their kernels in turn).
    hipStream_t s = one_stream ? stream : sts->s[c == 3 ? 2 : c];
    if (!one_stream && hipStreamWaitEvent(s, sts->start, 0) != hipSuccess) { err = "hipStreamWaitEvent failed"; return TKFMI_EHIP; }
    if (cp.bidi_cls[c]) {  // forward and backward of every pair in one launch
      CParams pc = prm;
      pc.ticket = flagv + 2 * c;
      const std::string kname = std::string("fbc_bidi_kernel<") + layout_name(pl) + "," + std::to_string(cp.W[c]) +
                                ">" + (c == 0 ? " long" : "");
      const int slot = tkfmi_ktime_start(kname.c_str(), cp.cls_pairs[c], cp.cls_cells[c], s, err);
      if (slot == -2) return TKFMI_EHIP;
      rc = launch_bidi(pl, cp.W[c], pc, pmat, ddesc + base[c][0], (int)cp.both[c].size(), cp.max_ly[c],
                       wpc[c][1], s, err);
      if (rc != TKFMI_OK) return rc;
      if (tkfmi_ktime_stop(slot, s, err) != TKFMI_OK) return TKFMI_EHIP;
    }
    for (int dir = 0; dir < (cp.bidi_cls[c] ? 0 : (fwd_only ? 1 : 2)); ++dir) {
      CParams pc = prm;
      pc.ticket = flagv + 2 * c + dir;
      const int n = (int)(dir ? cp.bwd[c].size() : cp.fwd[c].size());
      const std::string kname = std::string(dir ? "fbc_bwd_kernel<" : "fbc_fwd_kernel<") + layout_name(pl) +
                                "," + std::to_string(cp.W[c]) + ">" + (c == 0 ? " long" : "");
      const int slot = tkfmi_ktime_start(kname.c_str(), cp.cls_pairs[c], cp.cls_cells[c], s, err);
      if (slot == -2) return TKFMI_EHIP;
      rc = dispatch_chain(pl, cp.W[c], dir == 1, pc, pmat, ddesc + base[c][dir], n, cp.max_ly[c], wpc[c][dir], s, err);
      if (rc != TKFMI_OK) return rc;
      if (tkfmi_ktime_stop(slot, s, err) != TKFMI_OK) return TKFMI_EHIP;
    }
    if (one_stream) continue;
    if (hipEventRecord(sts->done[c], s) != hipSuccess) { err = "hipEventRecord failed"; return TKFMI_EHIP; }
    if (hipStreamWaitEvent(stream, sts->done[c], 0) != hipSuccess) { err = "hipStreamWaitEvent failed"; return TKFMI_EHIP; }
  }
  if (cp.bidi) {
    rc = launch_mark_failed(flagv + cp.fail_base, p->B, o->log_prob, stream, err);
    if (rc != TKFMI_OK) return rc;
  }
  if (o->status) {
    hipLaunchKernelGGL(raise_status_kernel, dim3(1), dim3(64), 0, stream, prm.abortf, o->status);
    if (hipGetLastError() != hipSuccess) { err = "status launch failed"; return TKFMI_EHIP; }
  }
  return TKFMI_OK;
}

}  // namespace

extern "C" int tkfmi_chain_config(int32_t long_w, int32_t main_w, double long_frac, int32_t long_one_launch) {
  if (long_w < 1 || long_w > 4 || main_w < 2 || main_w > 4 || !(long_frac >= 0.0))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_chain_config: waves must be 1..4 (main 2..4), long_frac >= 0");
  g_cfg.long_one = long_one_launch != 0;
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_cfg.w[0] = long_w;
  g_cfg.w[1] = main_w;
  g_cfg.long_frac = long_frac;
  g_tuned.clear();
  return TKFMI_OK;
}

extern "C" int tkfmi_chain_schedule(double cut, int32_t max_workgroups) {
  if (!(cut >= 0.0 && cut <= 1.0))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_chain_schedule: cut must be in [0, 1]");
  g_cut.store(cut);
  g_bidi_max_wg.store(max_workgroups < 0 ? -1 : max_workgroups);
  return TKFMI_OK;
}

namespace {
// Candidate (long_w, main_w, long_frac, long_one) settings timed by tkfmi_autotune; the
// current default for the batch's key is always timed first.
const ChainConfig kCandidates[] = {
    {{3, 3, 2, 1}, 0.7, 1},  {{3, 3, 2, 1}, 0.5, 1},  {{4, 3, 2, 1}, 0.6, 1},  {{3, 3, 2, 1}, 0.0, 0},
    {{3, 3, 2, 1}, 0.15, 0}, {{3, 3, 2, 1}, 0.3, 0},  {{4, 3, 2, 1}, 0.3, 0},  {{4, 4, 2, 1}, 0.3, 0},
    {{4, 4, 2, 1}, 0.0, 0},  {{3, 2, 2, 1}, 0.0, 0},  {{2, 2, 2, 1}, 0.0, 0}};
constexpr int kNCand = (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));

bool same_cfg(const ChainConfig& a, const ChainConfig& b) {
  return a.w[0] == b.w[0] && a.w[1] == b.w[1] && a.long_frac == b.long_frac && a.long_one == b.long_one;
}
}  // namespace

extern "C" size_t tkfmi_autotune_workspace_bytes(const tkfmi_problem* p, uint32_t flags) {
  Plan pl;
  if (!p || p->B < 0 || !p->state_types || !p->state_group || p->log_emit || !detect(p, pl))
    return tkfmi_fb2d_workspace_bytes(p, flags);  // a single configuration
  if (p->B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off)) return 0;
  size_t need = 0;
  for (int c = -1; c < kNCand; ++c) {
    const ChainPlan cp = plan_chain(p, pl, flags, c >= 0 ? &kCandidates[c] : nullptr);
    if (cp.ok) need = std::max(need, cp.total());
  }
  return need;
}

extern "C" int tkfmi_autotune(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                              void* stream_v, uint32_t flags, int32_t reps, tkfmi_tune_result* res) {
  tkfmi_clear_error();
  if (!res) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_autotune: result pointer missing");
  memset(res, 0, sizeof(*res));
  res->layout = -1;
  if (flags & (TKFMI_TIME_KERNELS | TKFMI_REUSE_DESCRIPTORS | TKFMI_GENERIC_KERNEL | TKFMI_CLASS_KERNEL))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_autotune: tunes the default (chained) path; drop the kernel, "
                                    "timing and reuse flags");
  Plan pl;
  if (!p || !p->state_types || !p->state_group || p->log_emit || !detect(p, pl))
    return TKFMI_OK;  // no tunable configuration for this layout (res->n = 0)
  if (p->B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_autotune: host length/offset arrays missing");
  if (reps < 1) reps = 3;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  const int key_l = (int)pl.lay, key_b = length_bin(max_seq_len(p));
  res->layout = key_l;
  res->length_bin = key_b;
  std::vector<ChainConfig> cand;
  cand.push_back(chain_config(key_l, key_b));  // the current choice first
  for (const ChainConfig& c : kCandidates)
    if (!same_cfg(c, cand[0])) cand.push_back(c);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return tkfmi_fail(TKFMI_EHIP, "tkfmi_autotune: hipEventCreate failed");
  int rc = TKFMI_OK;
  std::string err;
  const int nc = (int)std::min<size_t>(cand.size(), 16);
  for (int c = 0; c < nc; ++c) {
    res->long_w[c] = cand[c].w[0];
    res->main_w[c] = cand[c].w[1];
    res->long_frac[c] = cand[c].long_frac;
    res->long_one[c] = cand[c].long_one;
    res->ms[c] = -1.0;  // -1: not run (its workspace exceeds ws_bytes)
  }
  res->n = nc;
  // untimed warm-up on the first candidate (clocks, caches, first-use
  // set-up), then two round-robin passes; each candidate keeps its faster
  // pass, so no candidate is timed only on a cold GPU
  for (int r = 0; r < 2 * reps && rc == TKFMI_OK; ++r)
    if (ws && ws_bytes >= plan_chain(p, pl, flags, &cand[0]).total())
      rc = chain_run_impl(p, o, ws, ws_bytes, stream, flags, err, &cand[0]);
  for (int pass = 0; pass < 2 && rc == TKFMI_OK; ++pass) {
    for (int c = 0; c < nc && rc == TKFMI_OK; ++c) {
      const ChainPlan cp = plan_chain(p, pl, flags, &cand[c]);
      if (!cp.ok || !ws || ws_bytes < cp.total()) continue;
      rc = chain_run_impl(p, o, ws, ws_bytes, stream, flags, err, &cand[c]);  // warm-up
      if (rc == TKFMI_OK && hipEventRecord(e0, stream) != hipSuccess) rc = TKFMI_EHIP;
      for (int r = 0; r < reps && rc == TKFMI_OK; ++r)
        rc = chain_run_impl(p, o, ws, ws_bytes, stream, flags, err, &cand[c]);
      if (rc == TKFMI_OK && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = TKFMI_EHIP;
      float t = 0.f;
      if (rc == TKFMI_OK && hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
        const double ms = (double)t / reps;
        if (res->ms[c] < 0.0 || ms < res->ms[c]) res->ms[c] = ms;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // keep the first candidate unless another is > 3 % faster
  int best = 0;
  for (int c = 1; c < res->n; ++c)
    if (res->ms[c] > 0.0 && (res->ms[best] < 0.0 || res->ms[c] < res->ms[best])) best = c;
  if (res->ms[0] > 0.0 && !(res->ms[best] < 0.97 * res->ms[0])) best = 0;
  res->best = best;
  {
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    g_tuned[{key_l, key_b}] = cand[best];
  }
  if (rc != TKFMI_OK) return tkfmi_fail(rc, "tkfmi_autotune: " + (err.empty() ? std::string("HIP error") : err));
  return TKFMI_OK;
}
