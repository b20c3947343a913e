// tkfmi: 2D pair-HMM forward/backward + expected counts for gfx950 (MI355X).
//
// Reference algorithm: python/tkfmixdom/jax/dp/hmm.py (forward
// hmm.py:1186-1292, backward hmm.py:1295-1437, counts hmm.py:1930-1971).
// This file is an independent MI355X design, not a translation:
//
//  * One workgroup per pair (grid = B, pairs in decreasing-cost order so the
//    hardware dispatcher performs longest-first scheduling).
//  * A lattice cell is held by a group of GL lanes (GL = 16 for up to 15
//    emitting states, GL = 4 for up to 3); lane k of the group holds state
//    k, lane GL-1 carries the cell's binary exponent.  A wave therefore holds
//    64/GL cells, a workgroup H = W*64/GL consecutive lattice rows (a band).
//  * The band is swept by antidiagonals: at step t band-row R handles column
//    t - R.  When a row reaches the end of its band it wraps to the same row
//    of the next band, so the sweep never drains between bands; the band
//    boundary row goes through a small row buffer in global memory.
//  * Probability space with one exponent per cell: F(i,j,k) = m_k * 2^E(i,j).
//    The transition contraction is sum_s m_s P[s,k]: m_s is broadcast within
//    the lane group with DPP (row_newbcast / quad_perm), P[s,k] lives in the
//    lane's registers.  No exp/log in the inner loop.
//  * Forward pushes h = m . P to the type-shifted successor cells; the
//    backward pulls g (emission x B of the successors) and accumulates the
//    expected counts sum_c F[c,k] g_c[k'] in registers (15 per lane).
//  * The forward lattice (mantissas + exponent, GL doubles = 128 B per cell
//    for GL = 16) is stored to HBM row-major for the backward pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace {

constexpr int A_PAD = TKFMI_A_PAD;
constexpr int TAB_SZ = TKFMI_TAB_SZ;
constexpr int EXP_NONE = -(1 << 29);  // exponent of an all-zero cell
constexpr int ST_S = 0, ST_M = 1, ST_I = 2, ST_D = 3, ST_E = 4;
constexpr int MAX_SEQ_LDS = 32000;

struct PairDesc {
  int32_t pair;  // index in the caller's order
  int32_t Lx, Ly, pad;
  int64_t x_off, y_off;
  int64_t lat_off;   // doubles, into the lattice region
  int64_t row_off;   // doubles, into the row-buffer region
  int64_t post_off;  // doubles, into the posterior output
};

struct DevParams {
  int32_t S, nb, G, e_idx;
  int32_t n_trans, n_tab, flags, smem_seq_off;
  int32_t smem_rb_off;  // doubles: start of the LDS row buffer
  int32_t rb_cap;       // doubles available for the LDS row buffer
  int32_t smem_ecnt_off;  // doubles: emission-count tables (G*TAB_SZ)
  int32_t pad2;

  int32_t bidx[16];   // state index of lane k (-1: unused lane)
  int32_t btype[16];  // type of lane k (0: unused)
  int32_t bgrp[16];   // emission group of lane k
  const double* log_trans;
  const double* log_tab;
  const double* log_emit;  // dense mode: pair b's [i][j][state] table at desc.post_off
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  const PairDesc* desc;
  double* lattice;
  double* rowbuf;
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
};

// ---------------------------------------------------------------- DPP helpers
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = dpp_i32<CTRL>(__double2loint(v));
  const int hi = dpp_i32<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
// Value of lane SRC of this lane's group.
template <int GL, int SRC>
__device__ __forceinline__ double gbcast(double v) {
  if constexpr (GL == 16) {
    return dpp_f64<0x150 + SRC>(v);  // row_newbcast:SRC
  } else {
    return dpp_f64<SRC | (SRC << 2) | (SRC << 4) | (SRC << 6)>(v);  // quad_perm
  }
}
// Max over the lane group (every lane gets the result).
template <int GL>
__device__ __forceinline__ int gmax_i32(int v) {
  v = max(v, dpp_i32<0xB1>(v));  // quad_perm [1,0,3,2]
  v = max(v, dpp_i32<0x4E>(v));  // quad_perm [2,3,0,1]
  if constexpr (GL == 16) {
    v = max(v, dpp_i32<0x141>(v));  // row_half_mirror
    v = max(v, dpp_i32<0x140>(v));  // row_mirror
  }
  return v;
}
// Sum over the lane group; the butterfly pairs are symmetric so every lane
// obtains the bitwise-identical value.
template <int GL>
__device__ __forceinline__ double gsum_f64(double v) {
  v = v + dpp_f64<0xB1>(v);
  v = v + dpp_f64<0x4E>(v);
  if constexpr (GL == 16) {
    v = v + dpp_f64<0x141>(v);
    v = v + dpp_f64<0x140>(v);
  }
  return v;
}

__device__ __forceinline__ int frexp_exp(double v) {
  return __builtin_amdgcn_frexp_exp(v);
}

// Workgroup barrier that orders LDS only: it waits for LDS operations but not
// for outstanding global loads/stores, so the lattice stores and the
// prefetched lattice loads stay in flight across steps.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// acc += bcast(src, SRC) * c with the broadcast fused into the FP64 FMA
// (v_fmac_f64_dpp row_newbcast, 16-lane groups only).  The asm statements
// are volatile so they keep their order; NOP2 = true inserts the two wait
// states a DPP read needs after a VALU write of its source register, and is
// used on the first statement that reads a freshly computed source.
template <int SRC, bool NOP2>
__device__ __forceinline__ void fmac_bcast16(double& acc, double src, double c) {
  if constexpr (NOP2) {
    asm volatile("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "+v"(acc) : "v"(src), "v"(c), "n"(SRC));
  } else {
    asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "+v"(acc) : "v"(src), "v"(c), "n"(SRC));
  }
}

// h = sum_s bcast(m, s) * c[s]: two interleaved partial sums (even / odd s)
// combined at the end -- a fixed order, independent of the batch.
template <int GL, int NB, int... Is>
__device__ __forceinline__ double push_dot(double m, const double (&c)[NB],
                                           std::integer_sequence<int, Is...>) {
  double a0 = 0.0, a1 = 0.0;
  if constexpr (GL == 16) {
    ((Is % 2 == 0 ? fmac_bcast16<Is, Is == 0>(a0, m, c[Is])
                  : fmac_bcast16<Is, false>(a1, m, c[Is])), ...);
  } else {
    ((Is % 2 == 0 ? (void)(a0 = fma(gbcast<GL, Is>(m), c[Is], a0))
                  : (void)(a1 = fma(gbcast<GL, Is>(m), c[Is], a1))), ...);
  }
  return a0 + a1;
}
// Backward: b = sum_s prow[s] * bcast(g, s) (two partial sums) and
// acc[s] += bcast(g, s) * fs.
template <int GL, int NB, int... Is>
__device__ __forceinline__ double pull_dot_count(double g, const double (&prow)[NB],
                                                 double fs, double (&acc)[NB],
                                                 std::integer_sequence<int, Is...>) {
  double b0 = 0.0, b1 = 0.0;
  if constexpr (GL == 16) {
    ((Is % 2 == 0 ? fmac_bcast16<Is, Is == 0>(b0, g, prow[Is])
                  : fmac_bcast16<Is, false>(b1, g, prow[Is]),
      fmac_bcast16<Is, false>(acc[Is], g, fs)), ...);
  } else {
    double tmp;
    ((tmp = gbcast<GL, Is>(g),
      (Is % 2 == 0 ? (void)(b0 = fma(prow[Is], tmp, b0)) : (void)(b1 = fma(prow[Is], tmp, b1))),
      acc[Is] = fma(tmp, fs, acc[Is])), ...);
  }
  return b0 + b1;
}

// Normalise a per-lane value v * 2^e to the group: returns the mantissa and
// sets E to the group exponent (max |m| in [0.5, 1)).
template <int GL>
__device__ __forceinline__ double group_normalise(double v, int e, int& E) {
  const int le = (v != 0.0) ? e + frexp_exp(v) : EXP_NONE;
  int En = gmax_i32<GL>(le);
  if (En == EXP_NONE) En = 0;
  E = En;
  return (v != 0.0) ? ldexp(v, e - En) : 0.0;
}

// ------------------------------------------------------------------- kernel
// DENSE: emissions come from the caller's dense log table (the reference's
// log_emit_table) instead of the per-group tables; they are prefetched as
// log values PF steps ahead and exponentiated where used.
template <int GL, int NB, int W, bool DENSE>
__global__ __launch_bounds__(W * 64) void fb2d_kernel(DevParams P) {
  static_assert(NB <= GL - 1, "exponent lane must be free");
  constexpr int CPW = 64 / GL;
  constexpr int H = CPW * W;
  constexpr double LN2 = 0.69314718055994530942;
  using Seq = std::make_integer_sequence<int, NB>;

  extern __shared__ double smem[];
  const int G = P.G;
  double* s_tab = smem;                                   // G*TAB_SZ
  double* s_hbuf = s_tab + G * TAB_SZ;                    // 2*H*GL (also reduction)
  constexpr int HB = (2 * H * GL > W * GL * NB) ? 2 * H * GL : W * GL * NB;
  double* s_misc = s_hbuf + HB;                           // 8
  uint8_t* s_x = reinterpret_cast<uint8_t*>(smem) + P.smem_seq_off;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R = wave * CPW + lane / GL;  // band row held by this lane group
  const int k = lane % GL;               // state slot

  const PairDesc pd = P.desc[blockIdx.x];
  const int Lx = pd.Lx, Ly = pd.Ly, W1 = Ly + 1;
  const int S = P.S;
  uint8_t* s_y = s_x + ((Lx + 8) & ~7);

  // ---- stage emission tables (probability space) and sequences in LDS
  if constexpr (!DENSE) {
    const double* ltab = P.log_tab + (P.n_tab > 1 ? (size_t)pd.pair * G * TAB_SZ : 0);
    for (int q = tid; q < G * TAB_SZ; q += W * 64) s_tab[q] = exp(ltab[q]);
    for (int q = tid; q < Lx; q += W * 64) s_x[q] = min((int)P.x_cat[pd.x_off + q], A_PAD - 1);
    for (int q = tid; q < Ly; q += W * 64) s_y[q] = min((int)P.y_cat[pd.y_off + q], A_PAD - 1);
    if (tid == 0) {  // dummy characters for empty sequences (never used)
      if (Lx == 0) s_x[0] = 0;
      if (Ly == 0) s_y[0] = 0;
    }
  }
  const bool want_ecnt = !DENSE && !(P.flags & TKFMI_FORWARD_ONLY) && (P.flags & TKFMI_WANT_EMIT_COUNTS) &&
                         P.out_ecnt != nullptr;
  const double* dem = DENSE ? P.log_emit + pd.post_off : nullptr;  // [i][j][state]
  double* s_ecnt = smem + P.smem_ecnt_off;  // [G][TAB_SZ] expected emission counts
  if (want_ecnt)
    for (int q = tid; q < G * TAB_SZ; q += W * 64) s_ecnt[q] = 0.0;

  // ---- this lane's state, its transitions and emission addressing
  const double* lt = P.log_trans + (P.n_trans > 1 ? (size_t)pd.pair * S * S : 0);
  const int my_state = (k < NB) ? P.bidx[k] : -1;
  const int my_type = (my_state >= 0) ? P.btype[k] : 0;
  const int e_idx = P.e_idx;
  double pcol[NB];  // P[s -> me]
#pragma unroll
  for (int s = 0; s < NB; ++s)
    pcol[s] = (my_state >= 0 && P.bidx[s] >= 0) ? exp(lt[P.bidx[s] * S + my_state]) : 0.0;
  const double p_start = (my_state >= 0) ? exp(lt[my_state]) : 0.0;             // S -> me
  const double p_end = (my_state >= 0) ? exp(lt[my_state * S + e_idx]) : 0.0;    // me -> E
  const int ebase = P.bgrp[k < NB ? k : 0] * TAB_SZ +
                    (my_type == ST_I ? A_PAD * A_PAD : (my_type == ST_D ? A_PAD * A_PAD + A_PAD : 0));
  const int emx = (my_type == ST_M) ? A_PAD : (my_type == ST_D ? 1 : 0);
  const int emy = (my_type == ST_M || my_type == ST_I) ? 1 : 0;
  const bool want_lattice = !(P.flags & TKFMI_FORWARD_ONLY);
  const bool want_counts = want_lattice && !(P.flags & TKFMI_NO_COUNTS) && P.out_n != nullptr;
  const bool want_post = want_lattice && (P.flags & TKFMI_WANT_POSTERIORS) && P.out_post != nullptr;
  double* out_n = want_counts ? P.out_n + (size_t)pd.pair * S * S : nullptr;

  // Band bookkeeping.  Band-row R handles, at step t, position t - R of its
  // stream: band m = position / Pper, column index c = position % Pper.  The
  // period Pper >= H + PF + 1 guarantees that the row-buffer entry a band-row
  // 0 needs at step t was written (by row H-1 of the previous band) before
  // step t - PF, when it is prefetched.
  //
  // The row buffer lives in LDS when the row fits (rb_lds); otherwise in
  // global memory, where a full barrier (which also waits for global stores)
  // every QF steps and the longer period make each write visible before the
  // step that prefetches it.
  constexpr int PF = 6;   // prefetch distance in steps
  constexpr int QF = 32;  // full-barrier interval for a global row buffer
  const int nbands = (Lx + H) / H;  // ceil((Lx+1)/H)
  const int R_last = Lx - (nbands - 1) * H;
  const bool rb_lds = (W1 * GL) <= P.rb_cap;
  const int Pper = max(W1, rb_lds ? H + PF + 1 : H + PF + QF + 1);
  const int T_steps = (nbands - 1) * Pper + Ly + R_last + 1;
  double* lat = P.lattice + pd.lat_off;
  double* rowb = P.rowbuf + pd.row_off;
  double* s_rowb = smem + P.smem_rb_off;
  __syncthreads();

  // ======================================================== forward sweep
  {
    // position of the step being prefetched (t + PF) and of the current step
    int qm = 0, qc = -R;
    int m = 0, c = -R;
    double rb[PF];  // row-buffer values of band-row 0, PF steps ahead
    double de[PF];  // dense log emissions of this lane's state, PF steps ahead
    auto rb_load = [&](int m_, int c_) -> double {
      const bool need = (R == 0) && (m_ >= 1) && (m_ < nbands) && (c_ >= 0) && (c_ <= Ly) &&
                        (m_ * H <= Lx);
      if (!need) return 0.0;
      return rb_lds ? s_rowb[c_ * GL + k] : rowb[(size_t)c_ * GL + k];
    };
    auto e_load = [&](int m_, int c_) -> double {
      const int i_ = m_ * H + R;
      const bool v_ = (c_ >= 0) && (m_ < nbands) && (c_ <= Ly) && (i_ <= Lx) && my_state >= 0;
      return v_ ? dem[((size_t)i_ * W1 + c_) * S + my_state] : -1e30;
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      rb[u] = rb_load(qm, qc);
      if constexpr (DENSE) de[u] = e_load(qm, qc);
      if (++qc == Pper) { qc = 0; ++qm; }
    }
    double hI = 0.0; int eI = EXP_NONE;  // own h of the previous step
    double hM = 0.0; int eM = EXP_NONE;  // D-source of the previous step
    for (int t0 = 0; t0 < T_steps; t0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int t = t0 + u;
        if (t < T_steps) {
          const double rbv = rb[u];
          rb[u] = rb_load(qm, qc);
          double dlog = 0.0;
          if constexpr (DENSE) {
            dlog = de[u];
            de[u] = e_load(qm, qc);
          }
          if (++qc == Pper) { qc = 0; ++qm; }
          const int j = c;
          const int i = m * H + R;
          const bool valid = (c >= 0) && (m < nbands) && (j <= Ly) && (i <= Lx);
          // D-source: h of (i-1, j) -- previous step of band-row R-1, or the
          // row buffer for band-row 0
          const bool has_d = valid && i >= 1;
          double src = 0.0;
          if (has_d) src = (R > 0) ? s_hbuf[((((t - 1) & 1) * H) + (R - 1)) * GL + k] : rbv;
          const int src_e = (int)gbcast<GL, GL - 1>(src);
          const double dV = has_d ? src : 0.0;
          const int dE = has_d ? src_e : EXP_NONE;
          double sV; int sE; bool ok;
          if (my_type == ST_M) { sV = hM; sE = eM; ok = valid && i >= 1 && j >= 1; }
          else if (my_type == ST_I) { sV = hI; sE = eI; ok = valid && j >= 1; }
          else { sV = dV; sE = dE; ok = has_d && my_type == ST_D; }
          double em;
          if constexpr (DENSE) {
            em = exp(dlog);
          } else {
            const int xi = valid ? s_x[i >= 1 ? i - 1 : 0] : 0;
            const int yj = valid ? s_y[j >= 1 ? j - 1 : 0] : 0;
            em = s_tab[ebase + xi * emx + yj * emy];
          }
          const double v = ok ? sV * em : 0.0;
          int En;
          double mv = group_normalise<GL>(v, sE, En);
          const bool origin = valid && i == 0 && j == 0;
          if (origin) { mv = 0.0; En = 0; }
          if (want_lattice && valid)
            lat[((size_t)i * W1 + j) * GL + k] = (k == GL - 1) ? (double)En : mv;
          double h = push_dot<GL, NB>(mv, pcol, Seq{});
          if (origin) h = p_start;
          const double hout = (k == GL - 1) ? (double)En : h;
          s_hbuf[((t & 1) * H + R) * GL + k] = hout;
          if (valid && R == H - 1) {
            if (rb_lds) s_rowb[j * GL + k] = hout;
            else rowb[(size_t)j * GL + k] = hout;
          }
          if (valid && i == Lx && j == Ly) {  // corner: log P and the E column
            double tot = gsum_f64<GL>(mv * p_end);
            double lp;
            if (origin) {
              tot = exp(lt[e_idx]);  // empty pair: S -> E
              lp = lt[e_idx];
            } else {
              lp = log(tot) + (double)En * LN2;
            }
            const int ep = (tot > 0.0) ? En + frexp_exp(tot) : 0;
            const double mp = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
            if (k == 0) {
              P.out_logP[pd.pair] = (tot > 0.0) ? lp : -1e30;
              s_misc[0] = mp;
              s_misc[1] = (double)ep;
            }
            if (want_counts && tot > 0.0) {
              if (origin) {
                if (k == 0) out_n[e_idx] = 1.0;
              } else if (my_state >= 0) {
                out_n[my_state * S + e_idx] = mv * p_end / tot;
              }
            }
          }
          hI = h; eI = En;
          hM = dV; eM = dE;
          if (++c == Pper) { c = 0; ++m; }
          if (!rb_lds && (t & (QF - 1)) == QF - 1) __syncthreads();
          else lds_barrier();
        }
      }
    }
  }
  if (!want_lattice) return;

  // ======================================================= backward sweep
  // Band-row R of backward band mb holds lattice row i = Lx - mb*H - R and
  // visits columns j = Ly - c, c = 0, 1, ...
  const double mP = s_misc[0];
  const int EP = (int)s_misc[1];
  if (!(mP > 0.0)) return;  // impossible pair: counts/posteriors stay zero
  const double inv_mP = 1.0 / mP;
  double prow[NB];  // P[me -> s]
#pragma unroll
  for (int s = 0; s < NB; ++s)
    prow[s] = (my_state >= 0 && P.bidx[s] >= 0) ? exp(lt[my_state * S + P.bidx[s]]) : 0.0;
  double cacc[NB];
#pragma unroll
  for (int s = 0; s < NB; ++s) cacc[s] = 0.0;
  double* post = want_post ? P.out_post + pd.post_off : nullptr;
  __syncthreads();
  {
    int qm = 0, qc = -R;
    int mb = 0, c = -R;
    double fr[PF];  // forward-lattice values, PF steps ahead
    double rb[PF];  // row-buffer values of band-row 0, PF steps ahead
    double de[PF];  // dense log emissions, PF steps ahead
    auto f_load = [&](int mb_, int c_) -> double {
      const int i_ = Lx - mb_ * H - R;
      const bool v_ = (c_ >= 0) && (mb_ < nbands) && (c_ <= Ly) && (i_ >= 0);
      return v_ ? lat[((size_t)i_ * W1 + (Ly - c_)) * GL + k] : 0.0;
    };
    auto e_load = [&](int mb_, int c_) -> double {
      const int i_ = Lx - mb_ * H - R;
      const bool v_ = (c_ >= 0) && (mb_ < nbands) && (c_ <= Ly) && (i_ >= 0) && my_state >= 0;
      return v_ ? dem[((size_t)i_ * W1 + (Ly - c_)) * S + my_state] : -1e30;
    };
    auto rb_load = [&](int mb_, int c_) -> double {
      const bool need = (R == 0) && (mb_ >= 1) && (mb_ < nbands) && (c_ >= 0) && (c_ <= Ly) &&
                        (Lx - mb_ * H >= 0);
      if (!need) return 0.0;
      return rb_lds ? s_rowb[(Ly - c_) * GL + k] : rowb[(size_t)(Ly - c_) * GL + k];
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      fr[u] = f_load(qm, qc);
      rb[u] = rb_load(qm, qc);
      if constexpr (DENSE) de[u] = e_load(qm, qc);
      if (++qc == Pper) { qc = 0; ++qm; }
    }
    double qI = 0.0; int eI = EXP_NONE;  // own q of the previous step
    double qM = 0.0; int eM = EXP_NONE;  // D-successor of the previous step
    for (int t0 = 0; t0 < T_steps; t0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int t = t0 + u;
        if (t < T_steps) {
          const double fv_in = fr[u];
          const double rbv = rb[u];
          fr[u] = f_load(qm, qc);
          rb[u] = rb_load(qm, qc);
          double dlog = 0.0;
          if constexpr (DENSE) {
            dlog = de[u];
            de[u] = e_load(qm, qc);
          }
          if (++qc == Pper) { qc = 0; ++qm; }
          const int j = Ly - c;
          const int i = Lx - mb * H - R;
          const bool valid = (c >= 0) && (mb < nbands) && (c <= Ly) && (i >= 0);
          // D-successor: q of (i+1, j)
          const bool has_d = valid && i < Lx;
          double src = 0.0;
          if (has_d) src = (R > 0) ? s_hbuf[((((t - 1) & 1) * H) + (R - 1)) * GL + k] : rbv;
          const int src_e = (int)gbcast<GL, GL - 1>(src);
          const double dV = has_d ? src : 0.0;
          const int dE = has_d ? src_e : EXP_NONE;
          double sV; int sE; bool ok;
          if (my_type == ST_M) { sV = qM; sE = eM; ok = valid && i < Lx && j < Ly; }
          else if (my_type == ST_I) { sV = qI; sE = eI; ok = valid && j < Ly; }
          else { sV = dV; sE = dE; ok = has_d && my_type == ST_D; }
          int Eg;
          const double mg = group_normalise<GL>(ok ? sV : 0.0, sE, Eg);
          // forward value of this cell (prefetched)
          const double fv = valid ? fv_in : 0.0;
          const int EF = (int)gbcast<GL, GL - 1>(fv);
          const double fs = (my_state >= 0 && fv != 0.0) ? ldexp(fv, EF + Eg - EP) * inv_mP : 0.0;
          double bsum = pull_dot_count<GL, NB>(mg, prow, fs, cacc, Seq{});
          const bool corner = valid && i == Lx && j == Ly;
          int EBraw = Eg;
          if (corner) { bsum = p_end; EBraw = 0; }
          int EB;
          const double mB = group_normalise<GL>(bsum, EBraw, EB);
          if (valid && i == 0 && j == 0) {  // start state: S -> body counts, post[0,0,S]
            const double scale = ldexp(inv_mP, Eg - EP);
            if (want_counts && my_state >= 0) out_n[my_state] = p_start * mg * scale;
            const double b00 = gsum_f64<GL>(p_start * mg);
            if (want_post && k == 0) post[0] = corner ? 1.0 : b00 * scale;
          }
          if (want_post && valid && my_state >= 0)
            post[((size_t)i * W1 + j) * S + my_state] = fv * ldexp(mB * inv_mP, EF + EB - EP);
          // q = e(i,j,k) * B(i,j,k) for the predecessors
          const bool eok = (my_type == ST_M) ? (i >= 1 && j >= 1)
                         : (my_type == ST_I) ? (j >= 1)
                         : (my_type == ST_D) ? (i >= 1) : false;
          double q;
          if constexpr (DENSE) {
            q = eok ? exp(dlog) * mB : 0.0;
          } else {
            const int xi = valid ? s_x[i >= 1 ? i - 1 : 0] : 0;
            const int yj = valid ? s_y[j >= 1 ? j - 1 : 0] : 0;
            q = eok ? s_tab[ebase + xi * emx + yj * emy] * mB : 0.0;
            if (want_ecnt && valid && eok && my_state >= 0)
              atomicAdd(s_ecnt + ebase + xi * emx + yj * emy, fv * ldexp(mB * inv_mP, EF + EB - EP));
          }
          const double qout = (k == GL - 1) ? (double)EB : q;
          s_hbuf[((t & 1) * H + R) * GL + k] = qout;
          if (valid && R == H - 1) {
            if (rb_lds) s_rowb[j * GL + k] = qout;
            else rowb[(size_t)j * GL + k] = qout;
          }
          qI = q; eI = EB;
          qM = dV; eM = dE;
          if (++c == Pper) { c = 0; ++mb; }
          if (!rb_lds && (t & (QF - 1)) == QF - 1) __syncthreads();
          else lds_barrier();
        }
      }
    }
  }
  if (want_ecnt) {
    __syncthreads();
    double* oe = P.out_ecnt + (size_t)pd.pair * G * TAB_SZ;
    for (int q = tid; q < G * TAB_SZ; q += W * 64) oe[q] = s_ecnt[q];
  }
  if (!want_counts) return;

  // ---- reduce the count accumulators over the band rows (fixed order)
#pragma unroll
  for (int s = 0; s < NB; ++s) {
    double v = cacc[s];
#pragma unroll
    for (int off = GL; off < 64; off <<= 1) v += __shfl_xor(v, off);
    cacc[s] = v;
  }
  double* s_red = s_hbuf;  // [W][GL][NB]
  if (lane < GL) {
#pragma unroll
    for (int s = 0; s < NB; ++s) s_red[(wave * GL + lane) * NB + s] = cacc[s];
  }
  __syncthreads();
  for (int q = tid; q < NB * NB; q += W * 64) {
    const int kk = q / NB, s = q % NB;
    const int a = P.bidx[kk], b = P.bidx[s];
    if (a < 0 || b < 0) continue;
    double acc = 0.0;
    for (int w = 0; w < W; ++w) acc += s_red[(w * GL + kk) * NB + s];
    const double pr = exp(lt[a * S + b]);
    out_n[a * S + b] = (pr > 0.0) ? pr * acc : 0.0;
  }
}

// ------------------------------------------------------------- host side
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct Plan {
  int GL = 0, NB = 0, W = 0, H = 0;
  int S = 0, nb = 0, e_idx = -1;
  int bidx[16], btype[16], bgrp[16];
  size_t desc_bytes = 0, lat_doubles = 0, row_doubles = 0;
  int max_lx = 0, max_ly = 0;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int make_plan(const tkfmi_problem* p, uint32_t flags, Plan& pl) {
  if (!p) return fail(TKFMI_EINVAL, "null problem");
  if (p->B < 0) return fail(TKFMI_EINVAL, "B < 0");
  if (p->S < 2 || p->S > 4096) return fail(TKFMI_EINVAL, "S out of range");
  if (!p->state_types || !p->state_group) return fail(TKFMI_EINVAL, "state_types/state_group missing");
  if (p->n_groups < 1 || p->n_groups > 32) return fail(TKFMI_EINVAL, "n_groups must be in 1..32");
  if (p->state_types[0] != ST_S) return fail(TKFMI_EINVAL, "state 0 must have type S");
  pl.S = p->S;
  pl.nb = 0;
  for (int s = 0; s < p->S; ++s) {
    const int ty = p->state_types[s];
    if (ty < ST_S || ty > ST_E) return fail(TKFMI_EINVAL, "bad state type");
    if (ty == ST_E && pl.e_idx < 0) pl.e_idx = s;
    if (ty == ST_M || ty == ST_I || ty == ST_D) {
      if (pl.nb >= 15) return fail(TKFMI_EINVAL, "more than 15 emitting states is not supported by this kernel");
      const int g = p->state_group[s];
      if (g < 0 || g >= p->n_groups) return fail(TKFMI_EINVAL, "state_group out of range");
      pl.bidx[pl.nb] = s;
      pl.btype[pl.nb] = ty;
      pl.bgrp[pl.nb] = g;
      ++pl.nb;
    }
  }
  if (pl.e_idx < 0) return fail(TKFMI_EINVAL, "no state of type E");
  if (pl.nb <= 3) { pl.GL = 4; pl.NB = 3; pl.W = 4; }
  else { pl.GL = 16; pl.NB = 15; pl.W = 8; }
  pl.H = pl.W * 64 / pl.GL;
  for (int s = pl.nb; s < 16; ++s) { pl.bidx[s] = -1; pl.btype[s] = 0; pl.bgrp[s] = 0; }
  if (p->n_trans != 1 && p->n_trans != p->B) return fail(TKFMI_EINVAL, "n_trans must be 1 or B");
  if (!p->log_emit && p->n_tab != 1 && p->n_tab != p->B) return fail(TKFMI_EINVAL, "n_tab must be 1 or B");
  if (p->B > 0 && (!p->h_x_len || !p->h_y_len || !p->h_x_off || !p->h_y_off))
    return fail(TKFMI_EINVAL, "host length/offset arrays missing");
  pl.lat_doubles = 0;
  pl.row_doubles = 0;
  for (int b = 0; b < p->B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0) return fail(TKFMI_EINVAL, "negative length");
    if (lx + ly > MAX_SEQ_LDS) return fail(TKFMI_EINVAL, "Lx + Ly exceeds 32000");
    pl.max_lx = std::max(pl.max_lx, lx);
    pl.max_ly = std::max(pl.max_ly, ly);
    if (!(flags & TKFMI_FORWARD_ONLY))
      pl.lat_doubles += (size_t)(lx + 1) * (ly + 1) * pl.GL;
    pl.row_doubles += (size_t)(ly + 1) * pl.GL;
  }
  pl.desc_bytes = align_up(sizeof(PairDesc) * (size_t)std::max(p->B, 1), 256);
  return TKFMI_OK;
}

size_t plan_ws_bytes(const Plan& pl) {
  return pl.desc_bytes + align_up(pl.lat_doubles * 8, 256) + align_up(pl.row_doubles * 8, 256);
}

// Pinned staging buffers for host -> device uploads (per-pair descriptors,
// statistic programs).  A buffer is reused only once the copy out of it has
// completed (its event).  The pool grows while every buffer is still in
// flight, up to kMaxStages per host thread; beyond that the call waits for
// the earliest upload.  Growing costs a hipHostMalloc (milliseconds), so a
// small pool, filled by a caller's first calls, keeps a steady loop of calls
// from stalling on allocations while the host runs ahead of the GPU.
struct Stage {
  void* buf = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
  uint64_t seq = 0;  // submission order of its last upload
};
constexpr int kMaxStages = 8;
thread_local std::vector<Stage> g_stages;
thread_local uint64_t g_stage_seq = 0;

template <int GL, int NB, int W, bool DENSE>
int launch(const Plan& pl, DevParams dp, int B, size_t smem, hipStream_t stream) {
  auto kern = fb2d_kernel<GL, NB, W, DENSE>;
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   160 * 1024);
  });
  if (attr_err != hipSuccess) return fail(TKFMI_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(attr_err));
  hipLaunchKernelGGL(kern, dim3(B), dim3(W * 64), smem, stream, dp);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(TKFMI_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  (void)pl;
  return TKFMI_OK;
}

}  // namespace

int tkfmi_fail(int code, const std::string& msg) { return fail(code, msg); }
void tkfmi_clear_error() { g_err.clear(); }

int tkfmi_stage_upload(void* dst, const void* src, size_t bytes, hipStream_t stream,
                       std::string& err) {
  if (tkfmi_noalloc_active()) {
    // from pageable host memory: the runtime stages the bytes before the call
    // returns, so `src` may be released afterwards; nothing is allocated here
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) { err = std::string("hipMemcpyAsync: ") + hipGetErrorString(e); return TKFMI_EHIP; }
    return TKFMI_OK;
  }
  Stage* st = nullptr;
  for (Stage& c : g_stages) {  // a finished buffer, large enough if possible
    if (c.pending) {
      const hipError_t q = hipEventQuery(c.ev);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) { err = std::string("hipEventQuery: ") + hipGetErrorString(q); return TKFMI_EHIP; }
      c.pending = false;
    }
    if (!st || (st->cap < bytes && c.cap >= bytes)) st = &c;
    if (st->cap >= bytes) break;
  }
  if (!st) {
    if ((int)g_stages.size() < kMaxStages) {
      g_stages.emplace_back();
      st = &g_stages.back();
    } else {  // every buffer in flight: wait for the earliest upload
      st = &g_stages.front();
      for (Stage& c : g_stages)
        if (c.seq < st->seq) st = &c;
      hipError_t e = hipEventSynchronize(st->ev);
      if (e != hipSuccess) { err = std::string("hipEventSynchronize: ") + hipGetErrorString(e); return TKFMI_EHIP; }
      st->pending = false;
    }
  }
  if (st->cap < bytes) {
    if (st->buf) (void)hipHostFree(st->buf);
    st->buf = nullptr;
    st->cap = 0;
    hipError_t e = hipHostMalloc(&st->buf, bytes, hipHostMallocDefault);
    if (e != hipSuccess) { err = std::string("hipHostMalloc: ") + hipGetErrorString(e); return TKFMI_EHIP; }
    st->cap = bytes;
  }
  if (!st->ev) {
    hipError_t e = hipEventCreateWithFlags(&st->ev, hipEventDisableTiming);
    if (e != hipSuccess) { err = std::string("hipEventCreate: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  }
  memcpy(st->buf, src, bytes);
  hipError_t e = hipMemcpyAsync(dst, st->buf, bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) { err = std::string("hipMemcpyAsync: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  e = hipEventRecord(st->ev, stream);
  if (e != hipSuccess) { err = std::string("hipEventRecord: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  st->pending = true;
  st->seq = ++g_stage_seq;
  return TKFMI_OK;
}

namespace {
std::atomic<size_t> g_ckpt_budget{(size_t)256 << 20};
std::atomic<int> g_ckpt_segment{0};  // antidiagonals per checkpoint segment; 0 = sqrt rule
}

size_t tkfmi_checkpoint_budget_bytes() { return g_ckpt_budget.load(); }
int tkfmi_checkpoint_segment_k() { return g_ckpt_segment.load(); }

extern "C" {

int tkfmi_checkpoint_budget(size_t bytes) {
  g_ckpt_budget.store(bytes);
  return TKFMI_OK;
}

int tkfmi_checkpoint_segment(int k) {
  if (k != 0 && k < 4) {
    return fail(TKFMI_EINVAL, "tkfmi_checkpoint_segment: k must be 0 (automatic) or at least 4");
  }
  g_ckpt_segment.store(k);
  return TKFMI_OK;
}

int tkfmi_abi_version(void) { return TKFMI_ABI_VERSION; }

const char* tkfmi_last_error(void) { return g_err.c_str(); }

size_t tkfmi_fb2d_workspace_bytes(const tkfmi_problem* prob, uint32_t flags) {
  size_t v2b = 0;
  // dense emissions run on the generic or the wide kernel only
  const bool dense = prob && prob->log_emit;
  if (!dense && !(flags & (TKFMI_GENERIC_KERNEL | TKFMI_CLASS_KERNEL)) && tkfmi_chain_workspace(prob, flags, &v2b))
    return v2b > 0 ? v2b : 1;
  if (!dense && !(flags & TKFMI_GENERIC_KERNEL) && tkfmi_v2_workspace(prob, flags, &v2b))
    return v2b > 0 ? v2b : 1;
  if (prob && tkfmi_wide_applies(prob))
    return tkfmi_band_applies(prob, flags) ? tkfmi_band_workspace(prob, flags) : tkfmi_wide_workspace(prob, flags);
  Plan pl;
  if (make_plan(prob, flags, pl) != TKFMI_OK) return 0;
  return plan_ws_bytes(pl);
}

int tkfmi_fb2d_batched(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws,
                       size_t ws_bytes, void* stream_v, uint32_t flags) {
  g_err.clear();
  tkfmi_ktime_begin_call((flags & TKFMI_TIME_KERNELS) != 0);
  const bool dense = p && p->log_emit;
  if (dense && (flags & TKFMI_WANT_EMIT_COUNTS) && !(flags & TKFMI_FORWARD_ONLY))
    return fail(TKFMI_EINVAL, "TKFMI_WANT_EMIT_COUNTS needs table emissions (with dense log_emit, "
                              "d logP / d log_emit is the posterior lattice)");
  if (p && tkfmi_wide_applies(p)) {
    if (!o || !o->log_prob) return fail(TKFMI_EINVAL, "outputs.log_prob missing");
    if ((flags & TKFMI_WANT_POSTERIORS) && !(flags & TKFMI_FORWARD_ONLY) && !o->posteriors)
      return fail(TKFMI_EINVAL, "TKFMI_WANT_POSTERIORS without a posterior buffer");
    if ((flags & TKFMI_WANT_EMIT_COUNTS) && !(flags & TKFMI_FORWARD_ONLY) && !o->emit_counts)
      return fail(TKFMI_EINVAL, "TKFMI_WANT_EMIT_COUNTS without an emission-count buffer");
    if (p->B > 0 && (!p->log_trans || (!dense && (!p->log_tab || !p->x_cat || !p->y_cat))))
      return fail(TKFMI_EINVAL, "device input pointer missing");
    std::string err;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream_v);
    const int rcw = tkfmi_band_applies(p, flags) ? tkfmi_band_run(p, o, ws, ws_bytes, st, flags, err)
                                                 : tkfmi_wide_run(p, o, ws, ws_bytes, st, flags, err);
    return rcw == TKFMI_OK ? TKFMI_OK : fail(rcw, err);
  }
  Plan pl;
  int rc = make_plan(p, flags, pl);
  if (rc != TKFMI_OK) return rc;
  if (!o || !o->log_prob) return fail(TKFMI_EINVAL, "outputs.log_prob missing");
  if ((flags & TKFMI_WANT_POSTERIORS) && !(flags & TKFMI_FORWARD_ONLY) && !o->posteriors)
    return fail(TKFMI_EINVAL, "TKFMI_WANT_POSTERIORS without a posterior buffer");
  if ((flags & TKFMI_WANT_EMIT_COUNTS) && !(flags & TKFMI_FORWARD_ONLY) && !o->emit_counts)
    return fail(TKFMI_EINVAL, "TKFMI_WANT_EMIT_COUNTS without an emission-count buffer");
  if (p->B == 0) return TKFMI_OK;
  if (!p->log_trans || (!dense && (!p->log_tab || !p->x_cat || !p->y_cat)))
    return fail(TKFMI_EINVAL, "device input pointer missing");
  size_t v2b = 0;
  if (!dense && !(flags & (TKFMI_GENERIC_KERNEL | TKFMI_CLASS_KERNEL)) && tkfmi_chain_workspace(p, flags, &v2b)) {
    std::string err;
    const int rc2 = tkfmi_chain_run(p, o, ws, ws_bytes, reinterpret_cast<hipStream_t>(stream_v), flags, err);
    return rc2 == TKFMI_OK ? TKFMI_OK : fail(rc2, err);
  }
  if (!dense && !(flags & TKFMI_GENERIC_KERNEL) && tkfmi_v2_workspace(p, flags, &v2b)) {
    std::string err;
    const int rc2 = tkfmi_v2_run(p, o, ws, ws_bytes, reinterpret_cast<hipStream_t>(stream_v), flags, err);
    return rc2 == TKFMI_OK ? TKFMI_OK : fail(rc2, err);
  }
  const size_t need = plan_ws_bytes(pl);
  if (!ws || ws_bytes < need) return fail(TKFMI_EWORKSPACE, "workspace too small: need " + std::to_string(need));
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);

  // descriptors in decreasing-cost order
  const int B = p->B;
  std::vector<int> order(B);
  for (int b = 0; b < B; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int64_t ca = (int64_t)(p->h_x_len[a] + 1) * (p->h_y_len[a] + 1);
    const int64_t cb = (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
    return ca > cb;
  });
  std::vector<int64_t> post_off(B), lat_off(B), row_off(B);
  int64_t acc_post = 0, acc_lat = 0, acc_row = 0;
  for (int b = 0; b < B; ++b) {  // caller-order offsets for outputs
    post_off[b] = acc_post;
    acc_post += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1) * pl.S;
    lat_off[b] = acc_lat;
    if (!(flags & TKFMI_FORWARD_ONLY)) acc_lat += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1) * pl.GL;
    row_off[b] = acc_row;
    acc_row += (int64_t)(p->h_y_len[b] + 1) * pl.GL;
  }
  const size_t dbytes = sizeof(PairDesc) * (size_t)B;
  std::vector<PairDesc> hd(B);
  for (int q = 0; q < B; ++q) {
    const int b = order[q];
    PairDesc d;
    d.pair = b;
    d.Lx = p->h_x_len[b];
    d.Ly = p->h_y_len[b];
    d.pad = 0;
    d.x_off = p->h_x_off[b];
    d.y_off = p->h_y_off[b];
    d.lat_off = lat_off[b];
    d.row_off = row_off[b];
    d.post_off = post_off[b];
    hd[q] = d;
  }
  char* wsb = static_cast<char*>(ws);
  PairDesc* ddesc = reinterpret_cast<PairDesc*>(wsb);
  double* lattice = reinterpret_cast<double*>(wsb + pl.desc_bytes);
  double* rowbuf = reinterpret_cast<double*>(wsb + pl.desc_bytes + align_up(pl.lat_doubles * 8, 256));
  {
    std::string err;
    const int rcu = (flags & TKFMI_REUSE_DESCRIPTORS) ? TKFMI_OK : tkfmi_stage_upload(ddesc, hd.data(), dbytes, stream, err);
    if (rcu != TKFMI_OK) return fail(rcu, err);
  }

  DevParams dp;
  memset(&dp, 0, sizeof(dp));
  dp.S = pl.S;
  dp.nb = pl.nb;
  dp.G = p->n_groups;
  dp.e_idx = pl.e_idx;
  dp.n_trans = p->n_trans;
  dp.n_tab = p->n_tab;
  dp.flags = (int32_t)flags;
  for (int s = 0; s < 16; ++s) { dp.bidx[s] = pl.bidx[s]; dp.btype[s] = pl.btype[s]; dp.bgrp[s] = pl.bgrp[s]; }
  dp.log_trans = p->log_trans;
  dp.log_tab = p->log_tab;
  dp.log_emit = p->log_emit;
  dp.x_cat = p->x_cat;
  dp.y_cat = p->y_cat;
  dp.desc = ddesc;
  dp.lattice = lattice;
  dp.rowbuf = rowbuf;
  dp.out_logP = o->log_prob;
  dp.out_n = (flags & TKFMI_NO_COUNTS) ? nullptr : o->n_trans;
  dp.out_post = o->posteriors;
  dp.out_ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) ? o->emit_counts : nullptr;
  const int H = pl.H;
  const int HB = std::max(2 * H * pl.GL, pl.W * pl.GL * pl.NB);
  // LDS layout (doubles): tables | exchange/reduction | misc | row buffer | sequences
  const size_t kLdsBudget = 160 * 1024;
  const bool ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) && !(flags & TKFMI_FORWARD_ONLY);
  const size_t fixed_d = (size_t)p->n_groups * TAB_SZ * (ecnt ? 2 : 1) + HB + 8;
  dp.smem_ecnt_off = (int32_t)((size_t)p->n_groups * TAB_SZ + HB + 8);
  const size_t seq_bytes = dense ? 16 : (size_t)((pl.max_lx + 8) & ~7) + (size_t)((pl.max_ly + 8) & ~7) + 16;
  if (fixed_d * 8 + seq_bytes > kLdsBudget)
    return fail(TKFMI_EINVAL, "LDS budget exceeded (sequences too long for the emission groups)");
  const size_t rb_want = (size_t)(pl.max_ly + 1) * pl.GL;
  const size_t rb_room = (kLdsBudget - fixed_d * 8 - seq_bytes) / 8;
  const size_t rb_cap = std::min(rb_want, rb_room);
  dp.smem_rb_off = (int32_t)fixed_d;
  dp.rb_cap = (int32_t)rb_cap;
  const size_t seq_off = (fixed_d + rb_cap) * 8;
  dp.smem_seq_off = (int32_t)seq_off;
  const size_t smem = seq_off + seq_bytes;
  int64_t cells = 0;
  for (int b = 0; b < B; ++b) cells += (int64_t)(p->h_x_len[b] + 1) * (p->h_y_len[b] + 1);
  std::string err;
  const int slot = tkfmi_ktime_start(pl.GL == 4 ? "fb2d_kernel<4,3,4>" : "fb2d_kernel<16,15,8>", B,
                                     cells, stream, err);
  if (slot == -2) return fail(TKFMI_EHIP, err);
  if (dense)
    rc = (pl.GL == 4) ? launch<4, 3, 4, true>(pl, dp, B, smem, stream) : launch<16, 15, 8, true>(pl, dp, B, smem, stream);
  else
    rc = (pl.GL == 4) ? launch<4, 3, 4, false>(pl, dp, B, smem, stream) : launch<16, 15, 8, false>(pl, dp, B, smem, stream);
  if (rc != TKFMI_OK) return rc;
  if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return fail(TKFMI_EHIP, err);
  return TKFMI_OK;
}

}  // extern "C"
