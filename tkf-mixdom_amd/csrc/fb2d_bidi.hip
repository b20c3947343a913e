// tkfmi: chained cell-per-lane 2D pair-HMM forward/backward + expected
// counts for gfx950 (MI355X) with a pair's forward and backward running at
// the same time, from opposite corners, in one launch.  Reference algorithm:
// python/tkfmixdom/jax/dp/hmm.py (forward hmm.py:1186-1292, backward
// hmm.py:1295-1437, posteriors and counts hmm.py:1930-1971).
//
// The per-cell arithmetic and the chains of workgroups are those of
// fb2d_chain.hip.  There a pair's backward starts when its whole forward has
// ended, so a call lasts at least the longest pair's Lx + Ly forward steps
// plus as many backward steps.  Here the two sweeps overlap:
//  * A pair's antidiagonals d = i + j are split at C.  The forward computes F
//    on every cell and stores it on cells d <= C; the backward computes B on
//    every cell and, on cells d > C, stores its per-cell vector of successor
//    values g[k] = e_k(succ_k) B(succ_k, k) (succ_k: the cell state k moves
//    into) in the same lattice slots (a lattice row of one band step holds one
//    antidiagonal, so F rows and g rows never share a cache line).
//  * Expected counts are taken by the side that reaches a cell second: the
//    backward on cells d <= C (F from the lattice, as the two-launch
//    backward), the forward on cells d > C (g from the lattice): transition
//    counts F (x) g on the count wave, emission mass of state k at succ_k(c)
//    as h(c)[k] g(c)[k], with h = F P the forward's own push vector.
//  * Both scale by 1 / P(x, y).  P is the mass of the transitions that cross
//    the cut: the sum over cells c on d = C of h(c) . g(c), plus the M-type
//    part on d = C - 1 (an M move skips an antidiagonal), plus the end
//    transitions when C >= Lx + Ly.  The backward leaves g on d = C and the M
//    part of g on d = C - 1 in a cut buffer; every forward workgroup adds its
//    cells' terms and publishes its partial; every workgroup of the pair sums
//    the partials in the same order, so all use the same P.
//  * A pair's 2 G workgroups wait on one another (the forward on the
//    backward's cut values, both on the partials), so they hold consecutive
//    tickets: a pair whose tickets are all taken has all its workgroups
//    running and completes, and the pair being ticketed waits only for the
//    slots those free.  Waits are bounded as in fb2d_chain.hip.
//  * Across XCDs, each side makes its stores visible with one agent-scope
//    release (a workgroup barrier, then one thread's release atomic) and the
//    other side acquires before reading them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <type_traits>

#include "fb2d_cell.h"
#include "fb2d_chain.h"
#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_v2 {

// v 2^e += x 2^ex for x >= 0; v is kept in [0.5, 1) (or 0).
__device__ __forceinline__ void xacc(double& v, int& e, double x, int ex) {
  if (!(x > 0.0)) return;
  if (!(v > 0.0)) { v = x; e = ex; }
  else if (ex > e) { v = ldexp(v, e - ex) + x; e = ex; }
  else v += ldexp(x, ex - e);
  const int k = frexp_exp(v);
  v = ldexp(v, -k);
  e += k;
}

// Whole workgroup: sums the per-thread partials cv 2^ce into this forward
// group's slot (s_red: 2 NTHR doubles of LDS that nothing else uses now) and
// counts the slot in the pair's counter with an agent-scope release, which
// also publishes every store the workgroup made before (its lattice rows).
template <int NTHR>
__device__ void publish_partial(double cv, int ce, double* s_red, double* slot, int32_t* cnt) {
  s_red[2 * threadIdx.x] = cv;
  s_red[2 * threadIdx.x + 1] = (double)ce;
  wg_release_agent();  // the group's lattice rows go with its partial
  if (threadIdx.x < 64) {
    double v = 0.0;
    int e = 0;
    for (int q = threadIdx.x; q < NTHR; q += 64) xacc(v, e, s_red[2 * q], (int)s_red[2 * q + 1]);
    s_red[2 * threadIdx.x] = v;  // lane l read only entries l (mod 64)
    s_red[2 * threadIdx.x + 1] = (double)e;
    wave_lds_sync();
    if (threadIdx.x == 0) {
      double v2 = 0.0;
      int e2 = 0;
      for (int q = 0; q < 64; ++q) xacc(v2, e2, s_red[2 * q], (int)s_red[2 * q + 1]);
      slot[0] = v2;
      slot[1] = (double)e2;
      __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();  // s_red may be reused
}

// Whole wave: waits until every forward group of the pair has published its
// partial (acquire) and returns P = mP 2^EP, summed in group order.  False
// (and mP = 0) when the launch was aborted.
__device__ __forceinline__ bool wait_P(const CParams& Pr, const CDesc& pd, double& mP, int& EP) {
  const bool ok = wait_ge(Pr.flagv + pd.cnt, pd.gpair, Pr.abortf, Pr.spin_max);
  double v = 0.0;
  int e = 0;
  if (ok)
    for (int g = 0; g < pd.gpair; ++g) {
      const double* s = Pr.pslot + pd.pslot + 16 * g;
      xacc(v, e, s[0], (int)s[1]);
    }
  mP = v;
  EP = e;
  return ok;
}

// ============================================================== forward
// W recursion waves (one band each) plus one count wave (w == W).  Cells
// d <= C: the two-launch forward (F stored).  d = C - 1, C: terms of P from
// the cut buffer.  Cells d > C: g from the lattice, emission mass h (.) g at
// the successors and transition counts F (x) g.
template <class L, int W>
__device__ __forceinline__ void bidi_fwd(const CParams& Pr, const double* __restrict__ pmat, const CDesc& pd,
                                         double* smem, int* s_cflag) {
  constexpr int NTHR = (W + 1) * 64;
  constexpr int NB = L::NB, NG = L::NG;
  constexpr int nM = count_type<L, ST_M>();
  constexpr int nD = count_type<L, ST_D>();
  using LY = CLds<L, W>;
  constexpr int NX = LY::NX;
  constexpr int EIDX = L::EIDX;
  static_assert(NB <= SLOTS - 1, "exponent slot");

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int Lx = pd.Lx, Ly = pd.Ly, NW = pd.nw;
  const int S = Pr.S;
  const int BSTRIDE = Ly + 64;
  const int CS = Lx + 1;  // cut buffer stride

  double* s_tab = smem + LY::tab;
  double* s_ecnt = smem + LY::ecnt;
  double* s_ring = smem + LY::ring(true);
  double* s_xc = smem + LY::xc(true);
  double* s_stage = smem + LY::stage(true);
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + LY::seq_bytes(true);
  uint8_t* s_y = s_x + LY::xbytes;
  stage_inputs<L, W, NTHR>(Pr, pd, pmat + pd.pm_off, s_tab, s_x, s_y);
  if (threadIdx.x < W) s_cflag[threadIdx.x] = 0;
  const bool want_ecnt = (Pr.flags & TKFMI_WANT_EMIT_COUNTS) && Pr.out_ecnt != nullptr;
  if (want_ecnt)
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) s_ecnt[q] = 0.0;

  const double* __restrict__ pm = pmat + pd.pm_off;
  const double* __restrict__ Pf = pm;
  const double* p_start = smem + LY::misc;
  const double* p_end = p_start + NB;
  double prow[NB];  // prow[s] in lane l: P[s][l % 16] (0 past NB)
#pragma unroll
  for (int s = 0; s < NB; ++s) prow[s] = valu_copy((lane & 15) < NB ? pm[s * NB + (lane & 15)] : 0.0);

  const bool want_counts = !(Pr.flags & TKFMI_NO_COUNTS) && Pr.out_n != nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* lat = Pr.lattice + pd.lat_off;
  const double* cutb = Pr.cutbuf + pd.cut_off;
  const bool has_band = w < NW;
  const int band = pd.band0 + w;
  const int i_band = 64 * band + lane;
  const bool xin = (w == 0) && pd.fin >= 0;
  const bool xout = (w == NW - 1) && pd.fout >= 0;
  const bool counter = w == W;
  bool ok_run = true;
  const int T = pd.T;
  // Antidiagonal d is this workgroup's step t = d - 64 band0 (all waves).
  const int tC = pd.C - 64 * pd.band0;
  double mP = 0.0, inv_mP = 0.0;
  int EP = 0;
  d4 cacc = {0.0, 0.0, 0.0, 0.0}, cacc2 = {0.0, 0.0, 0.0, 0.0};
  double* stage_f = s_stage + (size_t)(counter ? 0 : w) * kStageDoubles;  // [64 cells][16] F-scaled
  double* stage_g = stage_f + kStageDoubles / 2;                          // [64 cells][16] g
  double accD[NG];   // D-type mass of row i + 1 (character x_i, fixed per lane)
  double runI[NG];   // I-type mass of the column passing down through the lanes
  // cut-step emission masses, binned once P is known: d = C - 1 (M) and d = C
  double kM1[NG], kM2[NG], kI2[NG], kD2[NG];
  int ke1, ke2;
#pragma unroll
  for (int gq = 0; gq < NG; ++gq) accD[gq] = 0.0;
  double cv = 0.0;  // this lane's terms of P: cv 2^ce
  int ce = 0;
  int put_cnt = 0, pub_cnt = 0;  // boundary entries written / published (lane 63 of the last wave)
  const int lane_last = min(63, Lx - 64 * band);  // the band's last row in the lattice
  __syncthreads();
  const int xi = s_x[64 * w + lane];
  const int xn = s_x[64 * w + lane + 1];  // x_i: the character of row i + 1

  // The meeting (whole workgroup): publish this group's partial of P, then
  // wait for the pair's other groups and bin the stashed cut-step masses.
  auto meet = [&](bool stashed) {
    if (xout && lane == 63 && put_cnt > pub_cnt) {  // entries the next group needs to reach the cut
      publish(Pr.flagv + pd.fout, put_cnt);
      pub_cnt = put_cnt;
    }
    publish_partial<NTHR>(cv, ce, s_stage, Pr.pslot + pd.pslot + 16 * pd.g, Pr.flagv + pd.cnt);
    if (ok_run) ok_run = wait_P(Pr, pd, mP, EP);
    inv_mP = (mP > 0.0) ? 1.0 / mP : 0.0;
    if (stashed && want_ecnt && has_band && i_band <= Lx) {
      // cells (i, j1) on d = C - 1 and (i, j2) on d = C; M mass to (x_i, y_j),
      // I mass to y_j (column j + 1), D mass to x_i (row i + 1)
      const int j1 = tC - 1 - 64 * w - lane, j2 = j1 + 1;
      const double ps1 = ldexp(inv_mP, ke1 - EP), ps2 = ldexp(inv_mP, ke2 - EP);
      const bool row_next = i_band < Lx;
#pragma unroll
      for (int gq = 0; gq < NG; ++gq) {
        if (row_next && j1 >= 0 && j1 < Ly && kM1[gq] != 0.0)
          atomicAdd(s_ecnt + gq * TAB_SZ + xn * A_PAD + s_y[j1], kM1[gq] * ps1);
        if (j2 >= 0 && j2 < Ly) {
          if (row_next && kM2[gq] != 0.0) atomicAdd(s_ecnt + gq * TAB_SZ + xn * A_PAD + s_y[j2], kM2[gq] * ps2);
          if (kI2[gq] != 0.0) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + s_y[j2], kI2[gq] * ps2);
        }
        if (row_next && kD2[gq] != 0.0) accD[gq] = fma(kD2[gq], ps2, accD[gq]);
      }
    }
  };

  if (tC < 0) {  // every cell of this group lies past the cut
    if (ok_run) ok_run = wait_ge(Pr.flagv + pd.bflag, 1, Pr.abortf, Pr.spin_max);
    meet(false);
  }

  double hp[NB];
  double rM_last[nM];
#pragma unroll
  for (int b = 0; b < NB; ++b) hp[b] = 0.0;
#pragma unroll
  for (int m = 0; m < nM; ++m) rM_last[m] = 0.0;
  int eH = EXP_NONE, eM_last = EXP_NONE;
  double emD[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) emD[g] = s_tab[g * TAB_SZ + A_PAD * A_PAD + A_PAD + xi];
  int yj_prev = 0;
  int y0_next = s_y[0];
  int fe = 0;  // boundary entries fetched into the chunk buffer
  const int t_on = 64 * w, t_off = 64 * w + 63 + Ly;

  // One step; phase 0: d < C - 1, 1: d = C - 1 or C, 2: d > C.  Each phase
  // runs as a loop of its own, so registers live in one phase only (the
  // cut-step stash, the g values, the count operands) do not add up.
  auto step = [&](auto phase, int t) {
    constexpr int PH = decltype(phase)::value;
    if (PH == 2 && counter && want_counts) count_step<W>(s_stage, NW, Ly, t, lane, s_cflag, cacc, cacc2, tC + 1);
    if (has_band && t >= t_on && t <= t_off) {
      int i = i_band;
      asm volatile("" : "+v"(i));  // keeps row predicates inside the loop (no SGPR masks)
      const int j = t - 64 * w - lane;
      const bool valid = j >= 0 && j <= Ly && i <= Lx;
      const bool cut1 = PH == 1 && t == tC - 1, cut2 = PH == 1 && t == tC;
      // ---- boundary row of the previous workgroup: fetch the next entries,
      // before the meeting only those this group needs to reach the cut
      if (xin && t <= Ly && t >= fe) {
        int e_end = min((t / CQ + 1) * CQ, Ly + 1);
        if (PH < 2) e_end = min(e_end, tC + 1);
        if (ok_run) ok_run = wait_ge(Pr.flagv + pd.fin, e_end, Pr.abortf, Pr.spin_max);
        fetch_chunk(s_xc + (size_t)(t % CQ) * NX, Pr.xring + pd.xin_off + (size_t)t * NX, (e_end - t) * NX);
        fe = e_end;
      }
      // ---- the backward's g of this cell: from the lattice past the cut (an
      // invalid cell reads the origin's slot, finite, and is scaled by 0),
      // from the cut buffer on d = C - 1 (M part) and d = C
      double gl[SLOTS];
      if constexpr (PH == 2) {
        const double* a = lat + (valid ? (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + 2 * lane : 0);
        load_slot<NB, false>(a, gl);
      } else if constexpr (PH == 1) {
        const bool okc = valid && (cut2 || (i < Lx && j < Ly));
        const double* a = cutb + (cut2 ? 16 * CS : 0) + (okc ? i : 0);
#pragma unroll
        for (int b = 0; b < NB; ++b) gl[b] = (okc && (cut2 || L::type(b) == ST_M)) ? a[b * CS] : 0.0;
        gl[SLOTS - 1] = okc ? a[(SLOTS - 1) * CS] : 0.0;
      }
      // ---- emissions at (i, j)
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
      // ---- new cell (no validity masks, as in the two-launch forward)
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
      // ---- store F on d <= C, where the backward counts
      if (PH < 2 && valid) {
        store_slot<NB, false>(lat + (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + 2 * lane, F, E);
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
      // ---- corner: log P and the end-transition counts
      if (valid && i == Lx && j == Ly) {
        double tot = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b) tot = fma(F[b], p_end[b], tot);
        double lp;
        if (origin) { tot = p_end[NB]; lp = log(tot); }
        else lp = log(tot) + (double)E * LN2;
        Pr.out_logP[pd.pair] = (tot > 0.0) ? lp : -1e30;
        if (want_counts && tot > 0.0) {
          if (origin) out_n[EIDX] = 1.0;
          else {
#pragma unroll
            for (int b = 0; b < NB; ++b) out_n[(b + L::SOFF) * S + EIDX] = F[b] * p_end[b] / tot;
          }
        }
        if (Lx + Ly <= pd.C) xacc(cv, ce, tot, E);  // the end transitions cross the cut
      }
      if constexpr (PH >= 1) {
        // ---- emission mass h(c)[k] g(c)[k] of state k at succ_k(c), per (group, type)
        double sM[NG], sI[NG], sD[NG];
#pragma unroll
        for (int gq = 0; gq < NG; ++gq) { sM[gq] = 0.0; sI[gq] = 0.0; sD[gq] = 0.0; }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ty = L::type(b), gq = L::group(b);
          const double m = h[b] * gl[b];
          if (ty == ST_M) sM[gq] += m;
          else if (ty == ST_I) sI[gq] += m;
          else sD[gq] += m;
        }
        const int Eg = (int)gl[SLOTS - 1];
        if constexpr (PH == 1) {  // terms of P; the masses wait for P in the stash
          double tot = 0.0;
#pragma unroll
          for (int gq = 0; gq < NG; ++gq) tot += sM[gq] + sI[gq] + sD[gq];
          if (valid) xacc(cv, ce, tot, E + Eg);
          // (an invalid cell's exponent is arbitrary: its stash stays 0)
          if (cut1 && valid) {
#pragma unroll
            for (int gq = 0; gq < NG; ++gq) kM1[gq] = sM[gq];
            ke1 = E + Eg;
          }
          if (cut2 && valid) {
#pragma unroll
            for (int gq = 0; gq < NG; ++gq) { kM2[gq] = sM[gq]; kI2[gq] = sI[gq]; kD2[gq] = sD[gq]; }
            ke2 = E + Eg;
          }
        } else {
          if (want_ecnt) {
            // M mass to (x_i, y_j), D mass to x_i (row i + 1, accumulated per
            // lane), I mass to y_j (column j + 1) through a running sum that
            // moves down the lanes with the column; the band's last row adds it.
            double cI[NG];
#pragma unroll
            for (int gq = 0; gq < NG; ++gq) cI[gq] = 0.0;
            if (valid) {
              const double pscale = ldexp(inv_mP, E + Eg - EP);
              const int yn = s_y[j];
#pragma unroll
              for (int gq = 0; gq < NG; ++gq) {
                if (i < Lx && j < Ly) atomicAdd(s_ecnt + gq * TAB_SZ + xn * A_PAD + yn, sM[gq] * pscale);
                if (j < Ly) cI[gq] = sI[gq] * pscale;
                if (i < Lx) accD[gq] = fma(sD[gq], pscale, accD[gq]);
              }
            }
#pragma unroll
            for (int gq = 0; gq < NG; ++gq) runI[gq] = dpp_f64<DPP_WAVE_SHR1>(runI[gq]) + cI[gq];
            if (lane == lane_last && j >= 0 && j < Ly) {
              const int yn = s_y[j];
#pragma unroll
              for (int gq = 0; gq < NG; ++gq) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + yn, runI[gq]);
            }
          }
          // ---- expected transition counts out of this cell: stage F and g
          if (want_counts) {
            const double fscale = valid ? ldexp(inv_mP, E + Eg - EP) : 0.0;
            int spins = 0;
            while (__hip_atomic_load(s_cflag + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < t) {
              __builtin_amdgcn_s_sleep(1);
              if (++spins > kSpinMax) { ok_run = false; break; }
            }
            int ln = lane;
            asm volatile("" : "+v"(ln));  // addresses computed here (hoisted, they would be spilled)
#pragma unroll
            for (int s = 0; s < 16; ++s) {
              const int o = ln * 16 + (s ^ (ln & 15));
              stage_f[o] = (s < NB) ? F[s < NB ? s : 0] * fscale : 0.0;
              stage_g[o] = (s < NB) ? gl[s < NB ? s : 0] : 0.0;
            }
          }
        }
      }
      // ---- publish this row's h for the band below (lane 63)
      if (lane == 63 && valid) {
        if (xout) {
          put_entry<L, NX>(Pr.xring + pd.xout_off + (size_t)j * NX, true, h, E);
          put_cnt = j + 1;
          if (((j + 1) % CQ) == 0 || j == Ly) {
            publish(Pr.flagv + pd.fout, j + 1);
            pub_cnt = j + 1;
          }
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
    if (PH == 1 && t == tC) meet(true);
  };
  const int t1 = max(0, min(tC - 1, T)), t2 = max(0, min(tC + 1, T));
  int t = 0;
  for (; t < t1; ++t) step(std::integral_constant<int, 0>(), t);
  if (t < t2) {
    // the backward's values of this group's rows (cut buffer, then the
    // lattice past the cut) are complete once its cut flag is up
    if (ok_run) ok_run = wait_ge(Pr.flagv + pd.bflag, 1, Pr.abortf, Pr.spin_max);
#pragma unroll
    for (int gq = 0; gq < NG; ++gq) { kM1[gq] = 0.0; kM2[gq] = 0.0; kI2[gq] = 0.0; kD2[gq] = 0.0; }
    ke1 = 0;
    ke2 = 0;
    for (; t < t2; ++t) step(std::integral_constant<int, 1>(), t);
  }
#pragma unroll
  for (int gq = 0; gq < NG; ++gq) runI[gq] = 0.0;
  for (; t < T; ++t) step(std::integral_constant<int, 2>(), t);
  if (tC == T) meet(true);  // the last step was d = C - 1: its M masses (into the next group) need P
  if (tC > T) {  // the cut lies past this group's cells: only its partial
    publish_partial<NTHR>(cv, ce, s_stage, Pr.pslot + pd.pslot + 16 * pd.g, Pr.flagv + pd.cnt);
  }
  if (counter && want_counts) count_step<W>(s_stage, NW, Ly, T, lane, s_cflag, cacc, cacc2, tC + 1);
  if (has_band && want_ecnt && i_band < Lx) {
#pragma unroll
    for (int gq = 0; gq < NG; ++gq)
      if (accD[gq] != 0.0) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + A_PAD + xn, accD[gq]);
  }
  __syncthreads();
  if (!ok_run) {  // the pair's log P becomes NaN after the launch (mark_failed_kernel)
    __hip_atomic_store(Pr.abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(Pr.flagv + Pr.fail_base + pd.pair, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (tC > T) return;  // no cell at or past the cut: nothing counted here
  if (want_ecnt) {  // layout group g -> caller group gmap[g]
    double* oe = Pr.out_ecnt + (size_t)pd.pair * Pr.G_tab * TAB_SZ;
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) {
      const int gq = q / TAB_SZ, o = q - gq * TAB_SZ;
      const double v = s_ecnt[q];
      if (v != 0.0) unsafeAtomicAdd(oe + Pr.gmap[gq] * TAB_SZ + o, v);
    }
  }
  if (!want_counts) return;
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

// ============================================================= backward
// W recursion waves (one band each) plus one count wave (w == W).  Cells
// d > C: the backward recursion only, g stored.  d = C: g and the M part of
// the next cell's g go to the cut buffer, then the meeting.  Cells d <= C:
// the two-launch backward (F from the lattice, counts).
template <class L, int W>
__device__ __forceinline__ void bidi_bwd(const CParams& Pr, const double* __restrict__ pmat, const CDesc& pd,
                                         double* smem, int* s_cflag) {
  constexpr int NTHR = (W + 1) * 64;
  constexpr int NB = L::NB, NG = L::NG;
  constexpr int nM = count_type<L, ST_M>();
  constexpr int nD = count_type<L, ST_D>();
  using LY = CLds<L, W>;
  constexpr int NX = LY::NX;

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int Lx = pd.Lx, Ly = pd.Ly, NW = pd.nw;
  const int S = Pr.S;
  const int BSTRIDE = Ly + 64;
  const int CS = Lx + 1;

  double* s_tab = smem + LY::tab;
  double* s_ecnt = smem + LY::ecnt;
  double* s_ring = smem + LY::ring(true);
  double* s_xc = smem + LY::xc(true);
  double* s_stage = smem + LY::stage(true);
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + LY::seq_bytes(true);
  uint8_t* s_y = s_x + LY::xbytes;
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
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* lat = Pr.lattice + pd.lat_off;
  double* cutb = Pr.cutbuf + pd.cut_off;
  const bool has_band = w < NW;
  const int band = pd.band0 + NW - 1 - w;  // wave 0 holds the bottom band
  const int i_band = 64 * band + lane;
  const bool xin = (w == 0) && pd.fin >= 0;
  const bool xout = (w == NW - 1) && pd.fout >= 0;
  bool ok_run = true;
  const bool counter = w == W;
  const int T = pd.T;
  // Antidiagonal d is this workgroup's step t = Ly + 64 (band0 + nw) - 1 - d.
  const int tC = Ly + 64 * (pd.band0 + NW) - 1 - pd.C;
  double mP = 0.0, inv_mP = 0.0;
  int EP = 0;
  d4 cacc = {0.0, 0.0, 0.0, 0.0}, cacc2 = {0.0, 0.0, 0.0, 0.0};
  double* stage_f = s_stage + (size_t)(counter ? 0 : w) * kStageDoubles;
  double* stage_g = stage_f + kStageDoubles / 2;
  double accD[NG];
  double runI[NG];
#pragma unroll
  for (int gq = 0; gq < NG; ++gq) accD[gq] = 0.0;
  int put_cnt = 0, pub_cnt = 0;  // boundary entries written / published (lane 0 of the last wave)
  __syncthreads();

  // The meeting (whole workgroup): raise this group's cut flag (an agent-scope
  // release of its g rows and cut buffer entries), then wait for P.
  auto meet = [&]() {
    if (xout && lane == 0 && put_cnt > pub_cnt) {  // entries the group above needs to reach the cut
      publish(Pr.flagv + pd.fout, put_cnt);
      pub_cnt = put_cnt;
    }
    wg_release_agent();
    if (threadIdx.x == 0) __hip_atomic_store(Pr.flagv + pd.bflag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (ok_run) ok_run = wait_P(Pr, pd, mP, EP);
    inv_mP = (mP > 0.0) ? 1.0 / mP : 0.0;
  };
  if (tC < 0) meet();  // every cell of this group lies before the cut

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
  int fe = 0;
  const int off = 64 * w + (63 - lane);  // position p = t - off, column j = Ly - p
  // Lattice slot of this lane's cell at position p.  An invalid cell reads
  // the origin's slot (cell (0, 0), always written by the forward, so finite);
  // its count operands are then scaled by 0.
  auto f_addr = [&](int p, int i) -> const double* {
    const bool v = has_band && p >= 0 && p <= Ly && i <= Lx;
    return lat + (v ? (((size_t)band * BSTRIDE + (Ly - p + lane)) * SLOTS) * 64 + 2 * lane : 0);
  };
  const int t_on = 64 * w, t_off = 64 * w + 63 + Ly;
  const bool corner_wave = has_band && band == (Lx + 64) / 64 - 1;
  const int t_corner = 64 * w + 63 - (Lx - 64 * band);

  // One step; phase 0: d > C (recursion, g stored), 1: d = C (cut values,
  // the meeting, then counts), 2: d < C (the two-launch backward step).
  // Each phase runs as a loop of its own (see the forward).
  auto step = [&](auto phase, int t) {
    constexpr int PH = decltype(phase)::value;
    if (PH >= 1 && counter && want_counts) count_step<W>(s_stage, NW, Ly, t, lane, s_cflag, cacc, cacc2, tC);
    const bool in_range = has_band && t >= t_on && t <= t_off;
    int i = i_band;
    asm volatile("" : "+v"(i));  // keeps row predicates inside the loop (no SGPR masks)
    const int p = t - off;
    const int j = Ly - p;
    const bool valid = p >= 0 && p <= Ly && i <= Lx;
    double fcur[SLOTS];
    double rD[nD > 0 ? nD : 1], rM[nM > 0 ? nM : 1];
    int eR = EXP_NONE;
    double g[NB];
    int Eg = 0;
    if (in_range) {
      // ---- boundary row of the next workgroup (below), up to the cut first
      // (before this step's lattice loads: loads complete in order, so the
      // chunk's wait would otherwise include them)
      if (xin && t <= Ly && t >= fe) {
        int e_end = min((t / CQ + 1) * CQ, Ly + 1);
        if (PH < 2) e_end = min(e_end, tC + 1);
        if (ok_run) ok_run = wait_ge(Pr.flagv + pd.fin, e_end, Pr.abortf, Pr.spin_max);
        fetch_chunk(s_xc + (size_t)(t % CQ) * NX, Pr.xring + pd.xin_off + (size_t)t * NX, (e_end - t) * NX);
        fe = e_end;
      }
      // The cell's forward values: loaded here, first used after the
      // transition contraction, which covers the load latency (at the cut:
      // after the meeting, which makes them visible).
      if constexpr (PH == 2) {
        const double* a = f_addr(p, i);
        load_slot<NB, false>(a, fcur);
      }
      // ---- receive the successor row's q (lane+1, or a ring for lane 63)
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
      Eg = max(cM, max(cI, cD));
      if (Eg == EXP_NONE) Eg = 0;
      const int shM = eM_last - Eg, shI = eQ - Eg, shD = eR - Eg;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ty = L::type(b);
        double v;
        if (ty == ST_M) v = okM ? ldexp(rM_last[rank_in_type<L, ST_M>(b)], shM) : 0.0;
        else if (ty == ST_I) v = okI ? ldexp(qp[b], shI) : 0.0;
        else v = okD ? ldexp(rD[rank_in_type<L, ST_D>(b)], shD) : 0.0;
        g[b] = v;
      }
      if constexpr (PH == 0) {  // past the cut: g for the forward's counts
        if (valid) {
          store_slot<NB, false>(lat + (((size_t)band * BSTRIDE + (j + lane)) * SLOTS) * 64 + 2 * lane, g, Eg);
        }
      }
      if constexpr (PH == 1) {  // the cut: g of this cell, the M part of the next cell's
        if (valid) {
#pragma unroll
          for (int b = 0; b < NB; ++b) cutb[(16 + b) * CS + i] = g[b];
          cutb[(16 + SLOTS - 1) * CS + i] = (double)Eg;
        }
        if (i <= Lx && j >= 1 && j - 1 <= Ly) {  // cell (i, j - 1); its M successor is (i + 1, j)
          const bool okn = i < Lx && j <= Ly;
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (L::type(b) == ST_M) cutb[b * CS + i] = okn ? rM[rank_in_type<L, ST_M>(b)] : 0.0;
          cutb[(SLOTS - 1) * CS + i] = okn ? (double)eR : 0.0;
        }
      }
    }
    if constexpr (PH == 1) {
      meet();
      if (in_range) {
        const double* a = f_addr(p, i);
        load_slot<NB, false>(a, fcur);
      }
    }
    if (in_range) {
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
      int EF = 0;
      if constexpr (PH >= 1) {
        // ---- counts: stage F-scaled and g, then MFMA (K = 4 cells per op)
        EF = (int)fcur[SLOTS - 1];
        const double fscale = valid ? ldexp(inv_mP, EF + Eg - EP) : 0.0;
        if (want_counts) {  // stage this step's cells (counted during the next step)
          int spins = 0;
          while (__hip_atomic_load(s_cflag + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < t) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinMax) { ok_run = false; break; }
          }
          int ln = lane;
          asm volatile("" : "+v"(ln));  // addresses computed here (hoisted, they would be spilled)
#pragma unroll
          for (int s = 0; s < 16; ++s) {
            const int o = ln * 16 + (s ^ (ln & 15));
            stage_f[o] = (s < NB) ? fcur[s < NB ? s : 0] * fscale : 0.0;
            stage_g[o] = (s < NB) ? g[s < NB ? s : 0] : 0.0;
          }
          if (valid && i == 0 && j == 0) {  // start state row: n[S, k]
            const double sc = ldexp(inv_mP, Eg - EP);
#pragma unroll
            for (int b = 0; b < NB; ++b) out_n[b + L::SOFF] = p_start[b] * g[b] * sc;
          }
        }
      }
      int yj = dpp_i32<DPP_WAVE_SHL1>(yj_prev);
      if (lane == 63) yj = y63_next;
      double emM[NG], emI[NG];
#pragma unroll
      for (int gq = 0; gq < NG; ++gq) {
        emM[gq] = s_tab[gq * TAB_SZ + xi * A_PAD + yj];
        emI[gq] = s_tab[gq * TAB_SZ + A_PAD * A_PAD + yj];
      }
      // ---- expected emission counts of cells d <= C (as the two-launch backward)
      if constexpr (PH >= 1) {
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
      }
      // ---- q = e(i,j,k) B(i,j,k) for the predecessor row
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
          put_cnt = p + 1;
          if (((p + 1) % CQ) == 0 || p == Ly) {
            publish(Pr.flagv + pd.fout, p + 1);
            pub_cnt = p + 1;
          }
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
  };
  const int tb1 = max(0, min(tC, T));
  int t = 0;
  for (; t < tb1; ++t) step(std::integral_constant<int, 0>(), t);
#pragma unroll
  for (int gq = 0; gq < NG; ++gq) runI[gq] = 0.0;
  if (t == tC && t < T) {
    step(std::integral_constant<int, 1>(), t);
    ++t;
  }
  for (; t < T; ++t) step(std::integral_constant<int, 2>(), t);
  if (tC >= T) {  // the cut lies before this group's cells: only its cut flag
    wg_release_agent();
    if (threadIdx.x == 0) __hip_atomic_store(Pr.flagv + pd.bflag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (counter && want_counts) count_step<W>(s_stage, NW, Ly, T, lane, s_cflag, cacc, cacc2, tC);
  if (has_band && want_ecnt && i_band >= 1 && i_band <= Lx) {
#pragma unroll
    for (int gq = 0; gq < NG; ++gq)
      if (accD[gq] != 0.0) atomicAdd(s_ecnt + gq * TAB_SZ + A_PAD * A_PAD + A_PAD + xi, accD[gq]);
  }
  __syncthreads();
  if (!ok_run) {  // the pair's log P becomes NaN after the launch (mark_failed_kernel)
    __hip_atomic_store(Pr.abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(Pr.flagv + Pr.fail_base + pd.pair, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (tC >= T) return;  // no cell before the cut: nothing counted here
  if (want_ecnt) {
    double* oe = Pr.out_ecnt + (size_t)pd.pair * Pr.G_tab * TAB_SZ;
    for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) {
      const int gq = q / TAB_SZ, o = q - gq * TAB_SZ;
      const double v = s_ecnt[q];
      if (v != 0.0) unsafeAtomicAdd(oe + Pr.gmap[gq] * TAB_SZ + o, v);
    }
  }
  if (!want_counts) return;
  double* s_red = s_stage;
  if (counter) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * v + (lane >> 4), col = lane & 15;
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

template <class L, int W>
__global__ __launch_bounds__((W + 1) * 64, 2) void fbc_bidi_kernel(CParams Pr, const double* __restrict__ pmat,
                                                                 const CDesc* __restrict__ descs) {
  extern __shared__ double smem[];
  __shared__ int s_cflag[W];  // count wave -> recursion wave w: stage read up to step s_cflag[w] - 1
  const CDesc pd = descs[take_ticket(Pr.ticket, Pr.abortf)];
  if (pd.dir == 0) bidi_fwd<L, W>(Pr, pmat, pd, smem, s_cflag);
  else bidi_bwd<L, W>(Pr, pmat, pd, smem, s_cflag);
}

// After every class of a one-launch call: the pairs one of whose waits
// timed out get a NaN log-likelihood (their other outputs are incomplete).
__global__ void mark_failed_kernel(const int32_t* __restrict__ fail, int B, double* __restrict__ out_logP) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x)
    if (fail[b]) out_logP[b] = __builtin_nan("");
}

int launch_mark_failed(const int32_t* fail, int B, double* out_logP, hipStream_t s, std::string& err) {
  if (B <= 0) return TKFMI_OK;
  hipLaunchKernelGGL(mark_failed_kernel, dim3(std::min(1024, (B + 255) / 256)), dim3(256), 0, s, fail, B, out_logP);
  if (hipGetLastError() != hipSuccess) { err = "mark_failed launch failed"; return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <class L, int W>
int launch_bidi_w(const CParams& prm, const double* pmat, const CDesc* d, int n, int max_ly, int wpc,
                  hipStream_t s, std::string& err) {
  const void* fn = (const void*)fbc_bidi_kernel<L, W>;
  size_t smem = CLds<L, W>::total(true, max_ly);
  if (wpc > 0) smem = std::max(smem, (size_t)(160 * 1024 / wpc - 512) & ~(size_t)255);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  hipLaunchKernelGGL((fbc_bidi_kernel<L, W>), dim3(n), dim3((W + 1) * 64), smem, s, prm, pmat, d);
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("fbc_bidi launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <class L>
int launch_bidi_l(int W, const CParams& prm, const double* pmat, const CDesc* d, int n, int max_ly, int wpc,
                  hipStream_t s, std::string& err) {
  switch (W) {
    case 4: return launch_bidi_w<L, 4>(prm, pmat, d, n, max_ly, wpc, s, err);
    case 3: return launch_bidi_w<L, 3>(prm, pmat, d, n, max_ly, wpc, s, err);
    case 2: return launch_bidi_w<L, 2>(prm, pmat, d, n, max_ly, wpc, s, err);
    default: return launch_bidi_w<L, 1>(prm, pmat, d, n, max_ly, wpc, s, err);
  }
}

int launch_bidi(const Plan& pl, int W, const CParams& prm, const double* pmat, const CDesc* d, int n,
                int max_ly, int wpc, hipStream_t s, std::string& err) {
  switch (pl.lay) {
    case Layout::Tkf: return launch_bidi_l<LTkf>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix11: return launch_bidi_l<LMix<1, 1>>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix21: return launch_bidi_l<LMix<2, 1>>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix31: return launch_bidi_l<LMix<3, 1>>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix12: return launch_bidi_l<LMix<1, 2>>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    case Layout::Mix13: return launch_bidi_l<LMix<1, 3>>(W, prm, pmat, d, n, max_ly, wpc, s, err);
    default: err = "layout not handled"; return TKFMI_EINVAL;
  }
}

}  // namespace tkfmi_v2
