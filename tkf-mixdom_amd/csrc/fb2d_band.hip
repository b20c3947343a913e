// tkfmi: banded forward/backward + expected counts for wide state spaces
// (16 to 160 emitting states, e.g. MixDom K=8 F=4: 2 + 160 states; the
// reference's trained K=3..8, F=2 models: 32..82 states) on gfx950.
// Reference algorithm: python/tkfmixdom/jax/dp/hmm.py (forward :1186-1292 and
// the row-scan variant :1440-1533 used for S > 50, backward :1295-1437 /
// :1536-1692, counts :1930-1971).
//
// Design:
//  * A pair's lattice rows are cut into bands of R rows (R = 16 for 96..160
//    padded states, 32 for 64, 64 for 32); each band is one workgroup of 8
//    waves, so a long pair spreads over many compute units.  Band g needs
//    the last row of band g - 1 (forward) or the first row of band g + 1
//    (backward): it receives it through a boundary buffer in global memory,
//    written with device-scope stores and published every CQ columns with a
//    progress counter (bounded waits).  Workgroups take their band from one
//    of eight ticket queues, the one of the XCD they run on; a pair's bands
//    sit in one queue, in dependency order, so a wait is only ever on a
//    workgroup that is already running, and a pair's transition block and
//    tables stay in that XCD's L2.
//  * Inside a band, step t handles cell (row r, column t - r) of every row:
//    the R cells of a step are independent.  Per step:
//      A (elementwise, one wave per cell, states over the lanes): the cell's
//        state vector from the values its predecessors pushed to it, times
//        the emissions (loaded one step ahead), normalised to one binary
//        exponent per cell with a single wave reduction; written to the HBM
//        lattice (forward) and to LDS;
//      B (matrix cores): the transition contraction of the step's cells as
//        (R x NBP) . (NBP x NBP) with v_mfma_f64_16x16x4f64.  The transition
//        block is held in registers for the whole sweep: each wave owns
//        column tiles over a slice of the contraction (KS slices), and each
//        output value is pushed straight to the one successor that uses it
//        (M-type states to (r+1, j+1), I-type to (r, j+1), D-type to
//        (r+1, j)) as KS partial sums in a two-slot LDS ring, summed in a
//        fixed order by the successor -- or, for the band's last row, to the
//        boundary buffer;
//      backward: B = g P^T the same way, the expected counts sum_c F_c g_c^T
//        as a second product (K = the step's cells) into 16 x 16 register
//        tiles held for the whole sweep, then a third elementwise phase
//        (posteriors, emission counts, q = e . B pushed to the predecessors).
//    Only the lattice and the boundary rows go to HBM.
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
constexpr int CQ = 16;                 // boundary columns per published chunk
constexpr int kYR = 256;               // LDS ring of descendant codes (chunks of 128)
constexpr int kQueues = 8;             // ticket queues, one per XCD
constexpr int kSpinMax = 1 << 22;      // bounded waits (each spin sleeps ~0.2 us + a load)
constexpr int kEcMaxG = 40;            // emission groups up to which counts are binned in LDS
constexpr int kEcM = A_PAD * A_PAD;    // M entries (x, y) of a group's table
constexpr double LN2 = 0.69314718055994530942;

typedef double d4 __attribute__((ext_vector_type(4)));

// Geometry per padded width NBP = 16 NT: rows per band R = 16 RT, the
// transition contraction's K (NBP / 4 chunks of 4) split into KS slices, so
// that NT x KS blocks (column tile, slice) spread evenly over the 4 SIMDs.
__host__ __device__ constexpr int band_rt(int NT) { return NT == 2 ? 2 : 1; }
// Workgroups per compute unit: two for up to 64 padded states (their
// per-step chains are latency-bound; a second workgroup fills the gaps),
// one above (LDS and registers).
__host__ __device__ constexpr int band_occ(int NT) { return NT <= 4 ? 2 : 1; }
// K slices of the forward's and the backward's products (the backward keeps
// one for 32 padded states: its LDS holds the emission-count tables too)
__host__ __device__ constexpr int band_ks(int NT, bool bwd) { return (NT == 8 || (bwd && NT == 2)) ? 1 : 2; }
template <int NT, bool BWD = false>
struct Geo {
  static constexpr int NBP = 16 * NT, NBPP = NBP + 1, KR = (NBP + 63) / 64;
  static constexpr int RT = band_rt(NT), R = 16 * RT, CPW = R / kWaves;
  // elementwise phases: LPC lanes per cell, CPI cells per wave instruction
  // (32 padded states: two cells per instruction, lanes 0-31 and 32-63),
  // NG groups of CPI cells per wave and step
  static constexpr int LPC = NBP <= 32 ? 32 : 64, CPI = 64 / LPC, NG = CPW / CPI;
  static constexpr int KS = band_ks(NT, BWD), KCH = NBP / 4, KB = KCH / KS;
  static constexpr int NBLK = NT * KS, BPW = (NBLK + kWaves - 1) / kWaves;
  static_assert(kWaves % KS == 0, "a wave's blocks share one slice (ks = w mod KS)");
  // fewer blocks than waves (32 padded states): wave w takes block w mod
  // NBLK for the row tiles rt = w / NBLK (mod RPW), so every wave has work
  static constexpr bool SPLIT = NBLK < kWaves;
  static constexpr int RPW = SPLIT ? kWaves / NBLK : 1;
  static_assert(!SPLIT || kWaves % NBLK == 0, "blocks divide the waves");
  static_assert(NG * CPI * kWaves == R, "rows w + kWaves (CPI cg + hc) are exactly 0..R-1");
  __host__ __device__ static constexpr int block(int w, int u) { return SPLIT ? w % NBLK : w + kWaves * u; }
  __host__ __device__ static constexpr int nblocks(int w) { return SPLIT ? 1 : (NBLK - w + kWaves - 1) / kWaves; }
  __host__ __device__ static constexpr int rt0(int w) { return SPLIT ? w / NBLK : 0; }
  static constexpr int NTILE = (NT * NT + kWaves - 1) / kWaves;  // count tiles per wave
};

struct BDesc {
  int32_t pair, Lx, Ly, i0;    // pair geometry; first row of the band
  int32_t nr, g, G, T;         // rows in the band, band index, bands of the pair, steps
  int32_t fin, fout;           // progress counters in / out (ints into the flag region; -1: none)
  int32_t ain, aout;           // backward: count-accumulator flags in (from band g+1) / out (to g-1)
  int64_t bin_off, bout_off;   // doubles into ws: boundary rows in / out ((Ly+1) x (NBP+1); -1: none)
  int64_t lat_off;             // doubles into ws: the band's lattice [T][R][NBP] by step and row, then
                               // its exponents [T][R] (a step's cells are contiguous)
  int64_t pm_off, tab_off;     // doubles into ws: transition block, probability-space tables
  int64_t x_off, y_off, post_off;
  int64_t acc_off;             // doubles into ws: the pair's count accumulator [NBP][NBP]
  int64_t mb_off;              // doubles into ws: the band's M-state posterior masses [T][R][nM]
  int64_t fbin_off;            // backward of a checkpointed band (lat_off < 0): the forward's
                               // boundary row g-1, from which the band's lattice is recomputed
};

struct BParams {
  int32_t S, NB, NBP, flags;
  int32_t e_idx, G, spin_max, nM;
  int32_t ec_mode, nslots, pad2, pad3;  // emission counts: 0 none, 1 binned (I, D in LDS; M masses
                                      // to a buffer), 2 atomics, 3 binned (I, D, M in LDS)
  const int32_t* mrank;  // [NBP] rank of slot k among the M-type slots (-1: other types)
  const int32_t* bidx;   // [NBP] state index of emitting slot k (padding: 0)
  const int32_t* btype;  // [NBP] type (0 for padding)
  const int32_t* bgrp;   // [NBP] emission group
  const int32_t* qoff;   // [kQueues + 1] first descriptor of each ticket queue
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  const double* log_emit;  // dense mode: pair b's [i][j][state] log table at desc.post_off
  const BDesc* desc;
  double* ws;
  int32_t* flagv;   // boundary progress counters
  int32_t* ticket;  // this launch's kQueues ticket counters
  int32_t* abortf;  // raised by a timed-out wait
  double* pinfo;    // [B][2]: mantissa and binary exponent of P(x, y)
  double* slots;    // checkpointed bands' recomputed lattices: nslots slots of slot_stride doubles
  int64_t slot_stride;
  int32_t* slotf;   // [nslots] slot in use
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
};

__host__ __device__ inline int64_t pm_stride(int NBP) { return 2 * (int64_t)NBP * NBP + 2 * NBP + 8; }

__device__ inline int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }

// One step of a max over lanes with DPP (row shifts within each row of 16
// lanes, row broadcasts of lanes 15 and 31); lanes whose DPP source is
// outside the row keep their own value, which max leaves unchanged.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_max_step(int v) {
  return max(v, __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xf, false));
}

// Binary exponent of the largest positive value over the lanes of a cell
// (the wave, or its half when CPI = 2; frexp convention), or EXP_NONE if no
// lane holds one.
template <int CPI>
__device__ __forceinline__ int cell_max_exp(double m) {
  int v = m > 0.0 ? frexp_exp(m) : EXP_NONE;
  v = dpp_max_step<0x111, 0xf>(v);  // row_shr:1
  v = dpp_max_step<0x112, 0xf>(v);  // row_shr:2
  v = dpp_max_step<0x114, 0xf>(v);  // row_shr:4
  v = dpp_max_step<0x118, 0xf>(v);  // row_shr:8
  v = dpp_max_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  if constexpr (CPI == 1) {
    v = dpp_max_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return __builtin_amdgcn_readlane(v, 63);
  } else {
    const int a = __builtin_amdgcn_readlane(v, 31), b = __builtin_amdgcn_readlane(v, 63);
    return (threadIdx.x & 32) ? b : a;
  }
}
// Sum over the lanes of a cell (the wave, or its half when CPI = 2).
template <int CPI>
__device__ __forceinline__ double cell_sum(double v) {
  for (int o = 32 / CPI; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ inline int emit_entry(int ty, int xi, int yj) {
  return ty == ST_M ? xi * A_PAD + yj : (ty == ST_I ? A_PAD * A_PAD + yj : A_PAD * A_PAD + A_PAD + xi);
}
__device__ inline bool emit_ok(int ty, int i, int j) {
  return ty == ST_M ? (i >= 1 && j >= 1) : (ty == ST_I ? j >= 1 : (ty == ST_D && i >= 1));
}
// Emission probability of a state of type ty and group g at cell (i, j):
// one load per lane, the entry chosen by selects (no branch per type, so
// that the loads of several cells stay in flight together).
__device__ inline double emit(const double* tab, int ty, int g, int xi, int yj, int i, int j) {
  return emit_ok(ty, i, j) ? tab[(size_t)g * TAB_SZ + emit_entry(ty, xi, yj)] : 0.0;
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
// The counter is read with relaxed loads: everything it publishes is read
// with device-scope loads at the coherence point (bnd_load), after this
// wave has seen the count, and the producer raised the count only once its
// stores had completed there.  An acquire would invalidate the XCD's L2
// (the transition blocks, tables and lattice lines of every workgroup on
// it) at each poll.
__device__ bool wait_ge(int* flag, int need, int* abortf, int spin_max) {
  int v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  int spins = 0;
  while (v < need) {
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      return false;
    if (++spins > spin_max) {
      __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
    v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  return true;
}

// The workgroup's band: the next ticket of its XCD's queue (hardware register
// XCC_ID), or of the next queue that still has bands.  Every queue lists its
// pairs' bands in dependency order, so the band a workgroup waits on holds an
// earlier ticket of the same queue and is already running.  The grid has one
// workgroup per band, so a workgroup always finds one; if none is left (a
// counter that was not zeroed) the abort flag is raised and the waves end
// without touching a descriptor.
__device__ __forceinline__ int take_band(int32_t* counters, const int32_t* qoff, int32_t* abortf) {
  __shared__ int s_ticket;
  if (threadIdx.x == 0) {
    const int xcc = __builtin_amdgcn_s_getreg((20 << 0) | (0 << 6) | (3 << 11)) & (kQueues - 1);
    int d = -1;
    for (int q = 0; q < kQueues && d < 0; ++q) {
      const int x = (xcc + q) & (kQueues - 1);
      const int n = qoff[x + 1] - qoff[x];
      if (n <= 0) continue;
      const int t = atomicAdd(counters + x, 1);
      if (t < n) d = qoff[x] + t;
    }
    s_ticket = d;
  }
  __syncthreads();
  const int t = __builtin_amdgcn_readfirstlane(s_ticket);
  if (t < 0) {
    if (threadIdx.x == 0) __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_endpgm();
  }
  return t;
}

// All of wave w's blocks (w + kWaves u, u < nb; they share the slice
// ks = w mod KS, hence the A operands) for row tile rt, interleaved so that
// the wave has nb independent accumulation chains.
template <class Gm, int NBW>
__device__ __forceinline__ void wave_gemm(const double* __restrict__ sA, int rt, int ks,
                                          const double (&fr)[NBW][Gm::KB], int nb, d4 (&acc)[NBW],
                                          int lane) {
  static_assert(Gm::KB % 4 == 0, "chunks in batches of 4");
  const double* ap = sA + (16 * rt + (lane & 15)) * Gm::NBPP + 4 * ks * Gm::KB + (lane >> 4);
#pragma unroll
  for (int u = 0; u < NBW; ++u) acc[u] = d4{0.0, 0.0, 0.0, 0.0};
  double a[4], an[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) a[q] = ap[4 * q];
#pragma unroll
  for (int cb = 0; cb < Gm::KB; cb += 4) {
    if (cb + 4 < Gm::KB) {
#pragma unroll
      for (int q = 0; q < 4; ++q) an[q] = ap[4 * (cb + 4 + q)];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < NBW; ++u)
        if (u < nb) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], fr[u][cb + q], acc[u], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = an[q];
  }
}

template <class Gm>
__device__ __forceinline__ d4 block_gemm_stream(const double* __restrict__ sA, int rt, int nt, int ks,
                                                const double* __restrict__ Bm, int lane) {
  static_assert(Gm::KB % 4 == 0, "chunks in batches of 4");
  const double* ap = sA + (16 * rt + (lane & 15)) * Gm::NBPP + 4 * ks * Gm::KB + (lane >> 4);
  const double* bp = Bm + (size_t)(4 * ks * Gm::KB + (lane >> 4)) * Gm::NBP + 16 * nt + (lane & 15);
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  double bn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) bn[q] = bp[(size_t)4 * q * Gm::NBP];
#pragma unroll 1
  for (int kb = 0; kb < Gm::KB; kb += 4) {
    double b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = bn[q];
    if (kb + 4 < Gm::KB) {
#pragma unroll
      for (int q = 0; q < 4; ++q) bn[q] = bp[(size_t)4 * (kb + 4 + q) * Gm::NBP];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ap[4 * (kb + q)], b[q], acc, 0, 0, 0);
  }
  return acc;
}

template <class Gm>
__device__ __forceinline__ void load_frag(const double* __restrict__ Bm, int b, int lane, double* fr) {
  const int nt = b / Gm::KS, ks = b - nt * Gm::KS;
  const double* bp = Bm + (size_t)(4 * ks * Gm::KB + (lane >> 4)) * Gm::NBP + 16 * nt + (lane & 15);
#pragma unroll
  for (int c = 0; c < Gm::KB; ++c) fr[c] = bp[(size_t)4 * c * Gm::NBP];
}

// Boundary entries [u0, u1) (entry u at column u, or Ly - u if reversed:
// NBP values and an exponent) into the LDS ring s_bin of CQ + 1 entries
// (slot u mod (CQ + 1)), which then holds the chunk and the last entry of
// the previous one.  Whole workgroup, two entries per wave, every load in
// flight at once.
template <int NBP>
__device__ __forceinline__ void fetch_bnd(double (*s_bin)[NBP + 1], const double* src, int u0, int u1,
                                          bool reversed, int Ly) {
  static_assert(CQ == 2 * kWaves, "two entries per wave");
  constexpr int K1 = (NBP + 1 + 63) / 64;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // w: wave-uniform (SGPR)
  double v[2][K1];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int u = u0 + w + kWaves * e;
    const double* src_e = src + (size_t)(reversed ? Ly - u : u) * (NBP + 1);
#pragma unroll
    for (int q = 0; q < K1; ++q) {
      const int k = lane + 64 * q;
      v[e][q] = (u < u1 && k <= NBP) ? bnd_load(src_e + k) : 0.0;
    }
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int u = u0 + w + kWaves * e;
#pragma unroll
    for (int q = 0; q < K1; ++q) {
      const int k = lane + 64 * q;
      if (u < u1 && k <= NBP) s_bin[u % (CQ + 1)][k] = v[e][q];
    }
  }
}

// Descendant codes ys[u] in an LDS ring of kYR: chunk [c0, c0 + 128) is
// loaded 64 steps before the first step that reads it.
__device__ __forceinline__ void load_y(uint8_t* s_y, const uint8_t* ys, int c0, int Ly) {
  const int u = c0 + (int)threadIdx.x;
  if (threadIdx.x < 128 && u < Ly) s_y[u & (kYR - 1)] = ys[u];
}

// ====================================================================== forward
// LDS of the forward sweep.
template <int NT>
struct FwdLds {
  using Gm = Geo<NT>;
  double s_pre[2][Gm::KS][Gm::R][Gm::NBPP];  // partial pushes to the cells of steps t (slot t&1), t+1
  double s_F[Gm::R][Gm::NBPP];               // this step's normalised forward values (GEMM operand)
  double s_last[Gm::KS][Gm::NBP];            // the last row's partial pushes for the band below
  double s_bin[CQ + 1][Gm::NBP + 1];         // the band above's last-row pushes (row 0's wave)
  int s_eM[2][Gm::R], s_eI[2][Gm::R], s_eD[2][Gm::R], s_E[Gm::R];
  int s_lastE, s_ok;                         // their exponent; cleared when a boundary wait gave up
  uint8_t s_y[kYR];
  double s_sink[64];                         // per-lane target of the matrix phase's discarded pushes
};

// One band's forward sweep.  RECOMP: the backward's recomputation of a
// checkpointed band -- its lattice goes to the workgroup's slot `lat`, the
// band above's pushes are read from the forward's boundary row `bin` (the
// forward launch has completed: no waits), and nothing is published.
template <int NT, bool DENSE, bool RECOMP>
__device__ __forceinline__ void fwd_sweep(const BParams& Pr, const BDesc& pd, unsigned char* smem, double* lat,
                                          const double* bin) {
  using Gm = Geo<NT>;
  constexpr int NBP = Gm::NBP, NBPP = Gm::NBPP, KR = Gm::KR, R = Gm::R;
  constexpr int KS = Gm::KS, KB = Gm::KB, NBLK = Gm::NBLK, BPW = Gm::BPW;
  FwdLds<NT>& L = *reinterpret_cast<FwdLds<NT>*>(smem);
  auto& s_pre = L.s_pre;
  auto& s_F = L.s_F;
  auto& s_last = L.s_last;
  auto& s_bin = L.s_bin;
  auto& s_eM = L.s_eM;
  auto& s_eI = L.s_eI;
  auto& s_eD = L.s_eD;
  auto& s_E = L.s_E;
  auto& s_lastE = L.s_lastE;
  auto& s_ok = L.s_ok;
  uint8_t* s_y = L.s_y;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // w: wave-uniform (SGPR)
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
  const bool want_lat = lat != nullptr;
  // the end transitions' counts come from the corner cell, also of a checkpointed pair (no lattice)
  const bool want_counts = !RECOMP && !(Pr.flags & (TKFMI_NO_COUNTS | TKFMI_FORWARD_ONLY)) && Pr.out_n != nullptr;
  double* latE = want_lat ? lat + (size_t)pd.T * R * NBP : nullptr;
  double* bout = (!RECOMP && pd.bout_off >= 0) ? Pr.ws + pd.bout_off : nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  constexpr int LPC = Gm::LPC, CPI = Gm::CPI, NG = Gm::NG;
  const int hc = (CPI == 2) ? (lane >> 5) : 0;  // the lane's cell within a group
  const int kl = lane & (LPC - 1);              // the lane's state slot (q = 0)
  int ty[KR], gr[KR], sx[KR];
#pragma unroll
  for (int q = 0; q < KR; ++q) {
    const int k = kl + 64 * q;
    ty[q] = (k < NB) ? Pr.btype[k] : 0;
    gr[q] = (k < NB) ? Pr.bgrp[k] : 0;
    sx[q] = (k < NB) ? Pr.bidx[k] : 0;
  }
  // the wave's transition blocks, held for the whole sweep, and the types of
  // the output columns it writes
  double fr[BPW][KB];
  int tkb[BPW];
  double psb[BPW];  // the origin's pushes p(S -> k) for the slice-0 blocks
  const int nbw = Gm::nblocks(w);  // the wave's blocks
#pragma unroll
  for (int u = 0; u < BPW; ++u) {
    const int b = Gm::block(w, u), k = 16 * (b / KS) + (lane & 15);
    if (b < NBLK) load_frag<Gm>(P, b, lane, fr[u]);
    tkb[u] = (b < NBLK && k < NB) ? Pr.btype[k] : 0;
    psb[u] = (b < NBLK && k < NB && b % KS == 0) ? p_start[k] : 0.0;
  }
  // the lane's row in group cg: r = w + kWaves (CPI cg + hc)
  int xr[NG];  // ancestor codes of the lane's rows
#pragma unroll
  for (int cg = 0; cg < NG; ++cg) {
    const int r = w + kWaves * (CPI * cg + hc), i = i0 + r;
    xr[cg] = (!DENSE && r < nr && i >= 1) ? (int)min((int)xs[i - 1], A_PAD - 1) : 0;
  }
  for (int q = threadIdx.x; q < 2 * KS * R * NBPP; q += kThr) (&s_pre[0][0][0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < R * NBPP; q += kThr) (&s_F[0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < KS * NBP; q += kThr) (&s_last[0][0])[q] = 0.0;
  if (threadIdx.x == 0) { s_ok = 1; s_lastE = 0; }
  if (threadIdx.x < 2 * R) {
    (&s_eM[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eI[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eD[0][0])[threadIdx.x] = EXP_NONE;
  }
  if constexpr (!DENSE) load_y(s_y, ys, 0, Ly);
  __syncthreads();
  // emissions (dense: log emissions) of the lane's cells at step tt
  double en[NG][KR];
  auto prefetch = [&](int tt) {
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int r = w + kWaves * (CPI * cg + hc), j = tt - r, i = i0 + r;
      const bool in = r < nr && j >= 0 && j <= Ly;
      const int yj = (!DENSE && in && j >= 1) ? (int)min((int)s_y[(j - 1) & (kYR - 1)], A_PAD - 1) : 0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        if constexpr (DENSE) en[cg][q] = (in && ty[q] != 0) ? dem[((size_t)i * W1 + j) * S + sx[q]] : 0.0;
        else en[cg][q] = in ? emit(tab, ty[q], gr[q], xr[cg], yj, i, j) : 0.0;
      }
    }
  };
  prefetch(0);
  bool ok_run = true;
  for (int t = 0; t < pd.T; ++t) {
    const int sl = t & 1, sl1 = sl ^ 1;
    ok_run = ok_run && s_ok;
    if constexpr (!DENSE) {
      if (((t + 64) & 127) == 0) load_y(s_y, ys, t + 64, Ly);
    }
    // ---- row 0's next chunk of the band above's pushes: wave 0 waits for
    // it, then the workgroup copies it to LDS
    if (bin != nullptr && (t % CQ) == 0 && t <= Ly) {
      const int u1 = min(t + CQ, Ly + 1);
      if (!RECOMP && w == 0 && ok_run) {
        ok_run = wait_ge(Pr.flagv + pd.fin, u1, Pr.abortf, Pr.spin_max);
        if (!ok_run && lane == 0) s_ok = 0;
      }
      __syncthreads();
      ok_run = ok_run && s_ok;
      if (ok_run) fetch_bnd<NBP>(s_bin, bin, t, u1, false, Ly);
      __syncthreads();
    }
    // ---- the last row's pushes of the previous step -> boundary (wave 7,
    // published after the matrix phase); a band whose input wait gave up
    // stops publishing, so that the bands below it give up too instead of
    // reading values computed from garbage
    const int jp = t - 1 - (nr - 1);
    const bool put = bout != nullptr && w == kWaves - 1 && ok_run && jp >= 0 && jp <= Ly;
    if (put) {
      double* e = bout + (size_t)jp * (NBP + 1);
      for (int k = lane; k < NBP; k += 64) {
        double h = s_last[0][k];
#pragma unroll
        for (int ks = 1; ks < KS; ++ks) h += s_last[ks][k];
        bnd_store(e + k, h);
      }
      if (lane == 0) bnd_store(e + NBP, (double)s_lastE);
    }
    // ---- A: the wave's cells of this step (CPI per instruction), in stages
    // across the groups: F' = v e 2^(e_type - E0) (E0: the largest
    // predecessor exponent), the cell's exponent from the largest F', then
    // F normalised
    const int tb0 = t % (CQ + 1), tb1 = tb0 == 0 ? CQ : tb0 - 1;  // ring entries t, t - 1
    double Fp[NG][KR];
    int E0c[NG], ex[NG];
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int r = w + kWaves * (CPI * cg + hc);
      const int j = t - r, i = i0 + r;
      const bool in = r < nr && j >= 0 && j <= Ly;
      // row 0 (the band above's entries) lies in wave 0's first group only,
      // where j = t: the entries j - 1 and j of the ring, a scalar index
      const bool topg = cg == 0 && w == 0 && bin != nullptr;  // wave-uniform
      const bool top = topg && r == 0;
      const int rc = r;  // < R: rows w + kWaves (CPI cg + hc) cover 0..R-1
      const bool okM = in && i >= 1 && j >= 1, okI = in && j >= 1, okD = in && i >= 1;
      // predecessor exponents: M from (i-1, j-1), I from (i, j-1), D from (i-1, j);
      // every read unconditional, the lane's value chosen by selects
      double bmv[KR], bdv[KR];
      int eMb = 0, eDb = 0;
      if (topg) {
        const double* bM = s_bin[tb1];  // entry j - 1
        const double* bD = s_bin[tb0];
        eMb = (int)bM[NBP];
        eDb = (int)bD[NBP];
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int k = min(kl + 64 * q, NBP - 1);
          bmv[q] = bM[k];
          bdv[q] = bD[k];
        }
      } else {
#pragma unroll
        for (int q = 0; q < KR; ++q) bmv[q] = bdv[q] = 0.0;
      }
      const int eMl = s_eM[sl][rc], eIl = s_eI[sl][rc], eDl = s_eD[sl][rc];
      const int eM = okM ? (top ? eMb : eMl) : EXP_NONE;
      const int eI = okI ? eIl : EXP_NONE;
      const int eD = okD ? (top ? eDb : eDl) : EXP_NONE;
      const int E0 = max(eM, max(eI, eD));
      E0c[cg] = E0;
      double m = 0.0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = min(kl + 64 * q, NBP - 1);
        double pre = s_pre[sl][0][rc][k];
#pragma unroll
        for (int ks = 1; ks < KS; ++ks) pre += s_pre[sl][ks][rc][k];
        const double bmv_ = bmv[q], bdv_ = bdv[q];
        const int tq = ty[q];
        double v = 0.0;  // selects, no branch per type
        v = (tq == ST_I && okI) ? pre : v;
        v = (tq == ST_M && okM) ? (top ? bmv_ : pre) : v;
        v = (tq == ST_D && okD) ? (top ? bdv_ : pre) : v;
        const int et = tq == ST_M ? eM : (tq == ST_I ? eI : eD);
        double e;
        if constexpr (DENSE) e = (v != 0.0) ? exp(en[cg][q]) : 0.0;
        else e = en[cg][q];
        const double f = (v != 0.0) ? ldexp(v * e, et - E0) : 0.0;
        Fp[cg][q] = f;
        m = fmax(m, f);
      }
      ex[cg] = cell_max_exp<CPI>(m);
    }
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int r = w + kWaves * (CPI * cg + hc);
      const int j = t - r, i = i0 + r;
      const bool in = r < nr && j >= 0 && j <= Ly;
      const int rc = r;
      const bool origin = (i == 0 && j == 0);
      const bool any = ex[cg] != EXP_NONE;
      const int sh = any ? ex[cg] : 0;
      const int E = (any && !origin) ? E0c[cg] + sh : 0;
      double F[KR];
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = kl + 64 * q;
        F[q] = (origin || !in) ? 0.0 : ldexp(Fp[cg][q], -sh);
        if (k < NBP) {
          s_F[rc][k] = F[q];
          if (want_lat && in) {  // written once, read once by the backward: non-temporal
            if constexpr (RECOMP) lat[((size_t)t * R + r) * NBP + k] = F[q];
            else __builtin_nontemporal_store(F[q], lat + ((size_t)t * R + r) * NBP + k);
          }
        }
      }
      if (kl == 0 && in) {
        s_E[r] = E;
        if (want_lat) {
          if constexpr (RECOMP) latE[(size_t)t * R + r] = (double)E;
          else __builtin_nontemporal_store((double)E, latE + (size_t)t * R + r);
        }
      }
      const bool corner = !RECOMP && in && i == Lx && j == Ly;
      if (__builtin_amdgcn_ballot_w64(corner)) {  // log P and the E column (the corner's cell)
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int k = kl + 64 * q;
          if (k < NB) part = fma(F[q], p_end[k], part);
        }
        double tot = cell_sum<CPI>(part);
        if (corner) {
          double lp;
          if (origin) { tot = p_se; lp = log(p_se); }
          else lp = log(tot) + (double)E * LN2;
          if (!ok_run) { tot = 0.0; lp = __builtin_nan(""); }
          if (kl == 0) {
            Pr.out_logP[pd.pair] = (tot > 0.0 || !ok_run) ? lp : -1e30;
            Pr.pinfo[2 * pd.pair] = (tot > 0.0) ? ldexp(tot, -frexp_exp(tot)) : 0.0;
            Pr.pinfo[2 * pd.pair + 1] = (tot > 0.0) ? (double)(E + frexp_exp(tot)) : 0.0;
          }
          if (want_counts && tot > 0.0) {
            if (origin) { if (kl == 0) out_n[Pr.e_idx] = 1.0; }
            else {
#pragma unroll
              for (int q = 0; q < KR; ++q) {
                const int k = kl + 64 * q;
                if (k < NB) out_n[Pr.bidx[k] * S + Pr.e_idx] = F[q] * p_end[k] / tot;
              }
            }
          }
        }
      }
    }
    if (t + 1 < pd.T) prefetch(t + 1);
    __syncthreads();
    // ---- B: h = F P for the step's cells, pushed to the successors as partial sums
    const bool origin_step = (i0 == 0 && t == 0);
    const int ks = w % KS;
#pragma unroll
    for (int rt = Gm::rt0(w); rt < Gm::RT; rt += Gm::RPW) {
      d4 acc[BPW];
      wave_gemm<Gm, BPW>(&s_F[0][0], rt, ks, fr, nbw, acc, lane);
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
        if (u >= nbw) break;
        const int k = 16 * (Gm::block(w, u) / KS) + (lane & 15);
        const int tk = tkb[u];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * rt + 4 * v + (lane >> 4);
          const int j = t - r;
          const bool skip = r >= nr || j < 0 || j > Ly || tk == 0;
          const double h = (origin_step && r == 0) ? psb[u] : acc[u][v];
          // one store, its address chosen by selects (no branch): I-type to
          // (r, j+1) (slot t+1), M-type to (r+1, j+1) (slot t+2), D-type to
          // (r+1, j) (slot t+1); the last row's M and D pushes to the
          // boundary row; a cell outside the band's rows or the pair's
          // columns, or a padding column, to the lane's sink
          const bool last = tk != ST_I && r + 1 >= nr;
          double* dst = skip ? &L.s_sink[lane]
                             : last ? &s_last[ks][k]
                                    : &s_pre[tk == ST_M ? sl : sl1][ks][tk == ST_I ? r : r + 1][k];
          *dst = h;
        }
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
          s_lastE = E;
        }
      }
    }
    if (put && (((jp + 1) % CQ) == 0 || jp == Ly)) publish(Pr.flagv + pd.fout, jp + 1);
    __syncthreads();
  }
  // ---- the last pushes -> boundary
  ok_run = ok_run && s_ok;
  if (bout && w == kWaves - 1 && ok_run) {
    const int jp = pd.T - 1 - (nr - 1);
    if (jp >= 0 && jp <= Ly) {
      double* e = bout + (size_t)jp * (NBP + 1);
      for (int k = lane; k < NBP; k += 64) {
        double h = s_last[0][k];
#pragma unroll
        for (int ks = 1; ks < KS; ++ks) h += s_last[ks][k];
        bnd_store(e + k, h);
      }
      if (lane == 0) bnd_store(e + NBP, (double)s_lastE);
      publish(Pr.flagv + pd.fout, jp + 1);
    }
  }
}

template <int NT, bool DENSE>
__global__ __launch_bounds__(kThr) __attribute__((amdgpu_waves_per_eu(2 * band_occ(NT), 8))) void band_fwd_kernel(
    BParams Pr) {
  __shared__ __align__(16) unsigned char smem[sizeof(FwdLds<NT>)];
  const BDesc pd = Pr.desc[take_band(Pr.ticket, Pr.qoff, Pr.abortf)];
  // checkpointed pairs (lat_off < 0) keep no lattice: the backward recomputes it
  double* lat = (!(Pr.flags & TKFMI_FORWARD_ONLY) && pd.lat_off >= 0) ? Pr.ws + pd.lat_off : nullptr;
  fwd_sweep<NT, DENSE, false>(Pr, pd, smem, lat, pd.bin_off >= 0 ? Pr.ws + pd.bin_off : nullptr);
}

// ===================================================================== backward
// The transition blocks are held in registers where they fit beside the
// count tiles (up to 96 padded states); above, they are read from L2 per
// step.
template <int NT>
__host__ __device__ constexpr bool bwd_breg() { return NT <= 6; }

// LDS of the backward sweep.
template <int NT>
struct BwdLds {
  using Gm = Geo<NT, true>;
  // (largest first: this order keeps the backward's address registers, and
  // with them its spills, lowest at 160 padded states)
  double s_B[Gm::KS][Gm::R][Gm::NBPP];  // partial sums of B = g P^T
  double s_pre[2][Gm::R][Gm::NBPP];   // q values pushed to the cells of steps t, t+1
  double s_bin[CQ + 1][Gm::NBP + 1];  // the band below's first-row pushes (row nr-1's wave)
  double s_G[Gm::R][Gm::NBPP];        // g of this step's cells (successor-pulled q)
  double s_FS[Gm::R][Gm::NBPP];       // their scaled forward values
  double s_last[Gm::NBP];             // the first row's pushes for the band above
  uint8_t s_y[kYR];
  int s_eM[2][Gm::R], s_eI[2][Gm::R], s_eD[2][Gm::R], s_Eg[Gm::R];
  int s_lastE, s_ok;                  // their exponent; cleared when a boundary wait gave up
  double s_sink[64];                  // per-lane target of the padding slots' pushes
};
// The backward kernel's static LDS: its own, or the forward sweep's for a
// checkpointed band's recomputation (the two are used one after the other).
template <int NT>
__host__ __device__ constexpr size_t bwd_smem() {
  return sizeof(FwdLds<NT>) > sizeof(BwdLds<NT>) ? sizeof(FwdLds<NT>) : sizeof(BwdLds<NT>);
}

// A free slot for a checkpointed band's lattice (the launch never has more
// such bands resident than slots), or -1 if the wait gave up.  The slot's
// previous holder may have run on another XCD, whose L2 is not coherent with
// this one's: it releases the slot with an agent-scope release (its dirty L2
// lines written back) and the taker acquires it (this XCD's stale copies
// dropped), so the lattice writes and reads below never mix two holders'
// lines.
__device__ int take_slot(int32_t* slotf, int nslots, int32_t* abortf, int spin_max) {
  __shared__ int s_slot;
  if (threadIdx.x == 0) {
    int got = -1;
    for (int spins = 0; got < 0 && spins <= spin_max; ++spins) {
      for (int q = 0; q < nslots && got < 0; ++q) {
        const int sl = (int)((blockIdx.x + q) % (unsigned)nslots);
        int expect = 0;
        if (__hip_atomic_load(slotf + sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
            __hip_atomic_compare_exchange_strong(slotf + sl, &expect, 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
          got = sl;
      }
      if (got < 0) __builtin_amdgcn_s_sleep(8);
    }
    if (got < 0) __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_slot = got;
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(s_slot);
}

// CK: the batch has checkpointed bands (a separate instantiation: the
// recomputation's registers would otherwise crowd the backward sweep's).
template <int NT, bool DENSE, bool CK>
__global__ __launch_bounds__(kThr) __attribute__((amdgpu_waves_per_eu(2 * band_occ(NT), 8))) void band_bwd_kernel(
    BParams Pr) {
  using Gm = Geo<NT, true>;
  constexpr int NBP = Gm::NBP, NBPP = Gm::NBPP, KR = Gm::KR, R = Gm::R;
  constexpr int KS = Gm::KS, KB = Gm::KB, NBLK = Gm::NBLK, BPW = Gm::BPW, NTILE = Gm::NTILE;
  constexpr bool BREG = bwd_breg<NT>();
  __shared__ __align__(16) unsigned char smem[bwd_smem<NT>()];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // w: wave-uniform (SGPR)
  const BDesc pd = Pr.desc[take_band(Pr.ticket, Pr.qoff, Pr.abortf)];
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
  const bool want_ecnt = !DENSE && Pr.ec_mode != 0 && Pr.out_ecnt != nullptr;
  const bool ec_bin = want_ecnt && (Pr.ec_mode == 1 || Pr.ec_mode == 3);
  const bool ec_mlds = want_ecnt && Pr.ec_mode == 3;
  // binned emission counts of this band: I-type entries [G][A_PAD] (by the
  // descendant residue), D-type [R][G] (by the row, whose ancestor residue
  // is fixed); M-type masses go to the band's buffer [T][R][nM]
  extern __shared__ double s_ec[];
  double* s_ecI = s_ec;
  double* s_ecD = s_ec + (size_t)Pr.G * A_PAD;
  double* s_ecM = s_ec + (size_t)Pr.G * (A_PAD + R);  // [G][A_PAD * A_PAD] (mode 3)
  double* mbuf = (ec_bin && !ec_mlds) ? Pr.ws + pd.mb_off : nullptr;
  const double* bin = pd.bin_off >= 0 ? Pr.ws + pd.bin_off : nullptr;
  double* bout = pd.bout_off >= 0 ? Pr.ws + pd.bout_off : nullptr;
  double* out_n = want_counts ? Pr.out_n + (size_t)pd.pair * S * S : nullptr;
  double* post = want_post ? Pr.out_post + pd.post_off : nullptr;
  double* ecnt = want_ecnt ? Pr.out_ecnt + (size_t)pd.pair * Pr.G * TAB_SZ : nullptr;
  const double mP = Pr.pinfo[2 * pd.pair];
  const int EP = (int)Pr.pinfo[2 * pd.pair + 1];
  if (!(mP > 0.0)) return;  // impossible or aborted pair: every band of it returns here
  const double inv_mP = 1.0 / mP;
  // the band's forward lattice: stored by the forward, or (checkpointed
  // band) recomputed here into a slot from the forward's boundary row
  const double* lat;
  int slot = -1;
  if (!CK || pd.lat_off >= 0) {
    lat = Pr.ws + pd.lat_off;
  } else {
    slot = take_slot(Pr.slotf, Pr.nslots, Pr.abortf, Pr.spin_max);
    if (slot < 0) {  // no slot (cannot happen with the planned count): give up on the pair
      if (threadIdx.x == 0) Pr.out_logP[pd.pair] = __builtin_nan("");
      return;
    }
    double* sl = Pr.slots + (size_t)slot * Pr.slot_stride;
    fwd_sweep<NT, DENSE, true>(Pr, pd, smem, sl, pd.fbin_off >= 0 ? Pr.ws + pd.fbin_off : nullptr);
    __syncthreads();
    lat = sl;
  }
  const double* latE = lat + (size_t)pd.T * R * NBP;
  auto release = [&] {  // the slot back to the pool, once the band has read it
    if (slot >= 0) {
      __syncthreads();
      if (threadIdx.x == 0) __hip_atomic_store(Pr.slotf + slot, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  };
  BwdLds<NT>& L = *reinterpret_cast<BwdLds<NT>*>(smem);
  auto& s_pre = L.s_pre;
  auto& s_G = L.s_G;
  auto& s_FS = L.s_FS;
  auto& s_B = L.s_B;
  auto& s_last = L.s_last;
  auto& s_bin = L.s_bin;
  auto& s_eM = L.s_eM;
  auto& s_eI = L.s_eI;
  auto& s_eD = L.s_eD;
  auto& s_Eg = L.s_Eg;
  auto& s_lastE = L.s_lastE;
  auto& s_ok = L.s_ok;
  uint8_t* s_y = L.s_y;
  constexpr int LPC = Gm::LPC, CPI = Gm::CPI, NG = Gm::NG;
  const int hc = (CPI == 2) ? (lane >> 5) : 0;  // the lane's cell within a group
  const int kl = lane & (LPC - 1);              // the lane's state slot (q = 0)
  int ty[KR], gr[KR], sx[KR];
#pragma unroll
  for (int q = 0; q < KR; ++q) {
    const int k = kl + 64 * q;
    ty[q] = (k < NB) ? Pr.btype[k] : 0;
    gr[q] = (k < NB) ? Pr.bgrp[k] : 0;
    sx[q] = (k < NB) ? Pr.bidx[k] : 0;
  }
  double fr[BREG ? BPW : 1][BREG ? KB : 1];
  if constexpr (BREG) {
#pragma unroll
    for (int u = 0; u < BPW; ++u)
      if (Gm::block(w, u) < NBLK) load_frag<Gm>(PT, Gm::block(w, u), lane, fr[u]);
  }
  // the lane's reversed row in group cg: rr = w + kWaves (CPI cg + hc)
  int xr[NG];  // ancestor codes of the lane's rows
#pragma unroll
  for (int cg = 0; cg < NG; ++cg) {
    const int rr = w + kWaves * (CPI * cg + hc), i = i0 + nr - 1 - rr;
    xr[cg] = (!DENSE && rr < nr && i >= 1) ? (int)min((int)xs[i - 1], A_PAD - 1) : 0;
  }
  for (int q = threadIdx.x; q < 2 * R * NBPP; q += kThr) (&s_pre[0][0][0])[q] = 0.0;
  for (int q = threadIdx.x; q < R * NBPP; q += kThr) {
    (&s_G[0][0])[q] = 0.0;
    (&s_FS[0][0])[q] = 0.0;
  }
  for (int q = threadIdx.x; q < NBP; q += kThr) s_last[q] = 0.0;
  if (ec_bin)
    for (int q = threadIdx.x; q < Pr.G * (A_PAD + R + (ec_mlds ? kEcM : 0)); q += kThr) s_ec[q] = 0.0;
  if (threadIdx.x == 0) { s_ok = 1; s_lastE = 0; }
  if (threadIdx.x < 2 * R) {
    (&s_eM[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eI[0][0])[threadIdx.x] = EXP_NONE;
    (&s_eD[0][0])[threadIdx.x] = EXP_NONE;
  }
  // Reversed coordinates: rr = nr - 1 - r (rr = 0 is the band's last row),
  // jj = Ly - j; step t handles (rr, jj = t - rr).  Successors of (i, j):
  // M (i+1, j+1) = (rr-1, jj-1), I (i, j+1) = (rr, jj-1), D (i+1, j) = (rr-1, jj).
  // The y ring holds ys[Ly - 1 - u] at u: cell (rr, jj) reads u = jj.
  auto load_yr = [&](int c0) {
    const int u = c0 + (int)threadIdx.x;
    if (threadIdx.x < 128 && u < Ly) s_y[u & (kYR - 1)] = ys[Ly - 1 - u];
  };
  if constexpr (!DENSE) load_yr(0);
  d4 cacc[NTILE];
#pragma unroll
  for (int q = 0; q < NTILE; ++q) cacc[q] = d4{0.0, 0.0, 0.0, 0.0};
  __syncthreads();
  // forward values and exponent of the wave's cells at step tt (the
  // forward's step T - 1 - tt, row nr - 1 - rr)
  double lf[NG][KR];
  int lfE[NG];
  auto prefetch_lat = [&](int tt) {
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int rr = w + kWaves * (CPI * cg + hc), jj = tt - rr;
      const bool in = rr < nr && jj >= 0 && jj <= Ly;
      const size_t cell = (size_t)(pd.T - 1 - tt) * R + (nr - 1 - rr);
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = kl + 64 * q;
        lf[cg][q] = (in && k < NBP) ? __builtin_nontemporal_load(lat + cell * NBP + k) : 0.0;
      }
      lfE[cg] = in ? (int)__builtin_nontemporal_load(latE + cell) : 0;
    }
  };
  // emissions (dense: log emissions) of the lane's cells at step tt
  double en[NG][KR];
  auto prefetch_em = [&](int tt) {
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int rr = w + kWaves * (CPI * cg + hc), jj = tt - rr, i = i0 + nr - 1 - rr, j = Ly - jj;
      const bool in = rr < nr && jj >= 0 && jj <= Ly;
      const int yj = (!DENSE && in && j >= 1) ? (int)min((int)s_y[jj & (kYR - 1)], A_PAD - 1) : 0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        if constexpr (DENSE) en[cg][q] = (in && ty[q] != 0) ? dem[((size_t)i * W1 + j) * S + sx[q]] : 0.0;
        else en[cg][q] = in ? emit(tab, ty[q], gr[q], xr[cg], yj, i, j) : 0.0;
      }
    }
  };
  prefetch_lat(0);
  prefetch_em(0);
  bool ok_run = true;
  for (int t = 0; t < pd.T; ++t) {
    const int sl = t & 1, sl1 = sl ^ 1;
    ok_run = ok_run && s_ok;
    if constexpr (!DENSE) {
      if (((t + 64) & 127) == 0) load_yr(t + 64);
    }
    // ---- the bottom row's next chunk of the band below's pushes: wave 0
    // waits for it, then the workgroup copies it to LDS
    if (bin != nullptr && (t % CQ) == 0 && t <= Ly) {
      const int u1 = min(t + CQ, Ly + 1);
      if (w == 0 && ok_run) {
        ok_run = wait_ge(Pr.flagv + pd.fin, u1, Pr.abortf, Pr.spin_max);
        if (!ok_run && lane == 0) s_ok = 0;
      }
      __syncthreads();
      ok_run = ok_run && s_ok;
      if (ok_run) fetch_bnd<NBP>(s_bin, bin, t, u1, true, Ly);
      __syncthreads();
    }
    // ---- the first row's pushes of the previous step -> boundary (wave 7,
    // published after the matrix phase)
    const int jjp = t - 1 - (nr - 1);
    const bool put = bout != nullptr && w == kWaves - 1 && ok_run && jjp >= 0 && jjp <= Ly;
    if (put) {
      double* e = bout + (size_t)(Ly - jjp) * (NBP + 1);
      for (int k = lane; k < NBP; k += 64) bnd_store(e + k, s_last[k]);
      if (lane == 0) bnd_store(e + NBP, (double)s_lastE);
    }
    // ---- A: g and the scaled forward values of this step's cells
    const int tb0 = t % (CQ + 1), tb1 = tb0 == 0 ? CQ : tb0 - 1;  // ring entries t, t - 1
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int rr = w + kWaves * (CPI * cg + hc);
      const int jj = t - rr, r = nr - 1 - rr, i = i0 + r, j = Ly - jj;
      const bool in = rr < nr && jj >= 0 && jj <= Ly;
      // the bottom row (the band below's entries) lies in wave 0's first
      // group only, where jj = t: the ring's entries jj - 1 and jj, a
      // scalar index; every read unconditional, values chosen by selects
      const bool botg = cg == 0 && w == 0 && bin != nullptr;  // wave-uniform
      const bool bot = botg && rr == 0;
      double bmv[KR], bdv[KR];
      int eMb = 0, eDb = 0;
      if (botg) {
        const double* bM = s_bin[tb1];  // column j + 1 (reversed entry jj - 1)
        const double* bD = s_bin[tb0];
        eMb = (int)bM[NBP];
        eDb = (int)bD[NBP];
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int kc = min(kl + 64 * q, NBP - 1);
          bmv[q] = bM[kc];
          bdv[q] = bD[kc];
        }
      } else {
#pragma unroll
        for (int q = 0; q < KR; ++q) bmv[q] = bdv[q] = 0.0;
      }
      const bool okM = in && i < Lx && j < Ly, okI = in && j < Ly, okD = in && i < Lx;
      const int eMl = s_eM[sl][rr], eIl = s_eI[sl][rr], eDl = s_eD[sl][rr];
      const int eM = okM ? (bot ? eMb : eMl) : EXP_NONE;
      const int eI = okI ? eIl : EXP_NONE;
      const int eD = okD ? (bot ? eDb : eDl) : EXP_NONE;
      int Eg = max(eM, max(eI, eD));
      if (Eg == EXP_NONE) Eg = 0;
      const double fsc = ldexp(inv_mP, lfE[cg] + Eg - EP);
      double g[KR];
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = kl + 64 * q;
        const int kc = min(k, NBP - 1);
        const double pre = s_pre[sl][rr][kc], bmv_ = bmv[q], bdv_ = bdv[q];
        const int tq = ty[q];
        double v = 0.0;  // selects, no branch per type
        v = (tq == ST_I && okI) ? pre : v;
        v = (tq == ST_M && okM) ? (bot ? bmv_ : pre) : v;
        v = (tq == ST_D && okD) ? (bot ? bdv_ : pre) : v;
        const int et = tq == ST_M ? eM : (tq == ST_I ? eI : eD);
        v = ldexp(v, et - Eg);
        g[q] = v;
        if (k < NBP) {
          s_G[rr][k] = v;
          s_FS[rr][k] = in ? lf[cg][q] * fsc : 0.0;
        }
      }
      if (kl == 0) s_Eg[rr] = Eg;
      const bool start = in && i == 0 && j == 0;
      if (__builtin_amdgcn_ballot_w64(start)) {  // start state: S -> body counts, posterior of S at (0, 0)
        const double scale = ldexp(inv_mP, Eg - EP);
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < KR; ++q) {
          const int k = kl + 64 * q;
          if (k < NB) {
            part = fma(p_start[k], g[q], part);
            if (want_counts && start) out_n[Pr.bidx[k]] = p_start[k] * g[q] * scale;
          }
        }
        part = cell_sum<CPI>(part);
        if (want_post && start && kl == 0) post[0] = (i == Lx && j == Ly) ? 1.0 : part * scale;
      }
    }
    if (t + 1 < pd.T) prefetch_lat(t + 1);
    __syncthreads();
    // ---- B: B = g P^T (partial sums per slice), counts += Fs^T g
    if constexpr (BREG) {
      const int ks = w % KS, nbw = Gm::nblocks(w);
#pragma unroll
      for (int rt = Gm::rt0(w); rt < Gm::RT; rt += Gm::RPW) {
        d4 acc[BPW];
        wave_gemm<Gm, BPW>(&s_G[0][0], rt, ks, fr, nbw, acc, lane);
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          if (u >= nbw) break;
          const int nt = Gm::block(w, u) / KS;
#pragma unroll
          for (int v = 0; v < 4; ++v) s_B[ks][16 * rt + 4 * v + (lane >> 4)][16 * nt + (lane & 15)] = acc[u][v];
        }
      }
    } else {
#pragma unroll 1
      for (int b = w; b < NBLK; b += kWaves) {
        const int nt = b / KS, ks = b - nt * KS;
#pragma unroll 1
        for (int rt = 0; rt < Gm::RT; ++rt) {
          const d4 acc = block_gemm_stream<Gm>(&s_G[0][0], rt, nt, ks, PT, lane);
#pragma unroll
          for (int v = 0; v < 4; ++v) s_B[ks][16 * rt + 4 * v + (lane >> 4)][16 * nt + (lane & 15)] = acc[v];
        }
      }
    }
    if (want_counts) {
#pragma unroll
      for (int kc = 0; kc < R / 4; ++kc) {
        const double* fa = &s_FS[4 * kc + (lane >> 4)][lane & 15];
        const double* ga = &s_G[4 * kc + (lane >> 4)][lane & 15];
#pragma unroll
        for (int q = 0; q < NTILE; ++q) {
          const int tile = w + kWaves * q;
          if (tile < NT * NT)
            cacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[16 * (tile / NT)], ga[16 * (tile % NT)], cacc[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    if (put && (((jjp + 1) % CQ) == 0 || jjp == Ly)) publish(Pr.flagv + pd.fout, jjp + 1);
    // ---- C: normalise B, posteriors, emission counts, q = e . B to the
    // predecessors; B's exponents of the lane's cells first, together
    double Bv[NG][KR];
    int ex[NG];
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int rr = w + kWaves * (CPI * cg + hc);
      const int jj = t - rr, i = i0 + nr - 1 - rr, j = Ly - jj;
      const bool in = rr < nr && jj >= 0 && jj <= Ly;
      const bool corner = (i == Lx && j == Ly);
      double mx = 0.0;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = kl + 64 * q;
        const int kc = min(k, NBP - 1);
        double bsum = s_B[0][rr][kc];
#pragma unroll
        for (int ks = 1; ks < KS; ++ks) bsum += s_B[ks][rr][kc];
        if (corner) bsum = (k < NB) ? p_end[kc] : 0.0;
        bsum = (in && k < NBP) ? bsum : 0.0;
        Bv[cg][q] = bsum;
        mx = fmax(mx, bsum);
      }
      ex[cg] = cell_max_exp<CPI>(mx);
    }
#pragma unroll
    for (int cg = 0; cg < NG; ++cg) {
      const int rr = w + kWaves * (CPI * cg + hc);
      const int jj = t - rr, r = nr - 1 - rr, i = i0 + r, j = Ly - jj;
      const bool in = rr < nr && jj >= 0 && jj <= Ly;
      const bool corner = (i == Lx && j == Ly);
      const int sh = (ex[cg] != EXP_NONE) ? ex[cg] : 0;
      const int EB = (corner ? 0 : s_Eg[rr]) + sh;
      const int yj = (!DENSE && in && j >= 1) ? (int)min((int)s_y[jj & (kYR - 1)], A_PAD - 1) : 0;
      const size_t cellS = ((size_t)i * W1 + j) * S;
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int k = kl + 64 * q;
        if (k >= NBP || !in) continue;
        const double b = ldexp(Bv[cg][q], -sh);
        double e = 0.0;
        if (k < NB) {
          if constexpr (DENSE) e = emit_ok(ty[q], i, j) ? exp(en[cg][q]) : 0.0;
          else e = en[cg][q];
        }
        const double qv = e * b;
        if (k < NB && (want_post || want_ecnt)) {
          // F B / P(x, y) = Fs (scaled by 2^(EF + Eg - EP) / mP) times B
          // before its normalisation (relative to 2^Eg; the corner: Eg = 0)
          const double pv = s_FS[rr][k] * Bv[cg][q];
          if (want_post) post[cellS + sx[q]] = pv;
          if (ec_bin) {
            const double pe = (e != 0.0) ? pv : 0.0;
            if (ec_mlds) {
              // every type binned in LDS: one atomic, its bin chosen by
              // selects (no branch per type)
              double* bin_ = ty[q] == ST_M ? s_ecM + gr[q] * kEcM + xr[cg] * A_PAD + yj
                             : ty[q] == ST_I ? s_ecI + gr[q] * A_PAD + yj
                                             : s_ecD + r * Pr.G + gr[q];
              if (pe != 0.0) atomicAdd(bin_, pe);
            } else if (ty[q] == ST_M) {
              if (i >= 1 && j >= 1) mbuf[((size_t)(pd.T - 1 - t) * R + r) * Pr.nM + Pr.mrank[k]] = pe;
            } else if (pe != 0.0) {
              if (ty[q] == ST_I) atomicAdd(s_ecI + gr[q] * A_PAD + yj, pe);
              else atomicAdd(s_ecD + r * Pr.G + gr[q], pe);
            }
          } else if (want_ecnt && pv != 0.0 && e != 0.0) {
            atomicAdd(ecnt + (size_t)gr[q] * TAB_SZ + emit_entry(ty[q], xr[cg], yj), pv);
          }
        }
        // push q: M-type to (i-1, j-1), I-type to (i, j-1), D-type to (i-1, j),
        // a padding slot's to the lane's sink (one store, its address chosen
        // by selects)
        const int tq = ty[q];
        const bool last = tq != ST_I && rr + 1 >= nr;
        double* dst = tq == 0 ? &L.s_sink[lane]
                      : last  ? &s_last[k]
                              : &s_pre[tq == ST_M ? sl : sl1][tq == ST_I ? rr : rr + 1][k];
        *dst = qv;
      }
      if (kl == 0 && in) {
        s_eI[sl1][rr] = EB;
        if (rr + 1 < nr) {
          s_eD[sl1][rr + 1] = EB;
          s_eM[sl][rr + 1] = EB;
        } else {
          s_lastE = EB;
        }
      }
    }
    if (t + 1 < pd.T) prefetch_em(t + 1);
    __syncthreads();
  }
  release();  // the lattice is read no more
  // ---- the last pushes -> boundary
  ok_run = ok_run && s_ok;
  if (bout && w == kWaves - 1 && ok_run) {
    const int jjp = pd.T - 1 - (nr - 1);
    if (jjp >= 0 && jjp <= Ly) {
      double* e = bout + (size_t)(Ly - jjp) * (NBP + 1);
      for (int k = lane; k < NBP; k += 64) bnd_store(e + k, s_last[k]);
      if (lane == 0) bnd_store(e + NBP, (double)s_lastE);
      publish(Pr.flagv + pd.fout, jjp + 1);
    }
  }
  if (ec_bin) {  // this band's binned I- and D-type (and in mode 3 M-type) emission counts
    __syncthreads();
    if (ec_mlds)
      for (int q = threadIdx.x; q < Pr.G * kEcM; q += kThr) {
        const double v = s_ecM[q];
        if (v != 0.0) atomicAdd(ecnt + (size_t)(q / kEcM) * TAB_SZ + q % kEcM, v);
      }
    for (int q = threadIdx.x; q < Pr.G * A_PAD; q += kThr) {
      const double v = s_ecI[q];
      if (v != 0.0) atomicAdd(ecnt + (size_t)(q / A_PAD) * TAB_SZ + kEcM + q % A_PAD, v);
    }
    for (int q = threadIdx.x; q < nr * Pr.G; q += kThr) {
      const double v = s_ecD[q];
      const int r = q / Pr.G, i = i0 + r;
      if (v != 0.0 && i >= 1)
        atomicAdd(ecnt + (size_t)(q % Pr.G) * TAB_SZ + kEcM + A_PAD + min((int)xs[i - 1], A_PAD - 1), v);
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

// M-type emission counts of one pair (one workgroup per pair): the
// backward's per-cell M masses binned by (group, ancestor residue,
// descendant residue) in LDS, then written to the pair's tables (their M
// entries; the I and D entries come from the bands' atomics).
struct EcPair {
  int32_t pair, Lx, Ly, G;
  int32_t g0, g1, pad0, pad1;    // this workgroup's bands [g0, g1) of the pair
  int64_t x_off, y_off, mb_off;  // mb_off: band g0's masses; band g's follow band g-1's
};
// Bands per reduction workgroup: about kEcChunk mass values each, so that a
// long pair's masses are read by several compute units at once.
constexpr int64_t kEcChunk = 2 << 20;

__global__ __launch_bounds__(512) void band_ecnt_m_kernel(const EcPair* __restrict__ ep, int R, int nM, int G,
                                                          const int32_t* __restrict__ mgrp,
                                                          const uint8_t* __restrict__ x_cat,
                                                          const uint8_t* __restrict__ y_cat,
                                                          const double* __restrict__ ws, const double* __restrict__ pinfo,
                                                          double* __restrict__ out_ecnt) {
  extern __shared__ double s_acc[];  // [G][A_PAD * A_PAD], then the pair's clamped codes x, y
  const EcPair e = ep[blockIdx.x];
  if (!(pinfo[2 * e.pair] > 0.0)) return;  // impossible pair: its backward wrote no masses
  for (int q = threadIdx.x; q < G * kEcM; q += blockDim.x) s_acc[q] = 0.0;
  // the codes in LDS: each mass needs its cell's two residues, and global
  // byte loads beside every mass load held this kernel to a third of HBM speed
  uint8_t* xs = reinterpret_cast<uint8_t*>(s_acc + G * kEcM);
  uint8_t* ys = xs + e.Lx;
  for (int q = threadIdx.x; q < e.Lx; q += blockDim.x) xs[q] = (uint8_t)min((int)x_cat[e.x_off + q], A_PAD - 1);
  for (int q = threadIdx.x; q < e.Ly; q += blockDim.x) ys[q] = (uint8_t)min((int)y_cat[e.y_off + q], A_PAD - 1);
  __syncthreads();
  const double* mb = ws + e.mb_off;
  // thread -> (cell slot, M rank): nMp = nM rounded up to a power of two,
  // 512 / nMp cells per pass; R is a power of two (16, 32)
  int nMp = 1;
  while (nMp < nM) nMp <<= 1;
  const int m = threadIdx.x & (nMp - 1), cs = threadIdx.x / nMp, cpp = blockDim.x / nMp;
  const int lr = __builtin_ctz(R);
  const int gm = (m < nM) ? mgrp[m] : 0;
  for (int g = e.g0; g < e.g1; ++g) {
    const int nr = min(R, e.Lx + 1 - g * R), T = e.Ly + nr, ncell = T * R;
    if (m < nM) {
      constexpr int U = 4;  // cells per thread and pass, their loads in flight together
      for (int base = cs; base < ncell; base += U * cpp) {
        double v[U];
        int xy[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int cell = base + u * cpp;
          const int r = cell & (R - 1), t = cell >> lr;
          const int i = g * R + r, j = t - r;
          const bool ok = cell < ncell && r < nr && j >= 1 && j <= e.Ly && i >= 1;
          v[u] = ok ? mb[cell * nM + m] : 0.0;
          xy[u] = ok ? (int)xs[i - 1] * A_PAD + (int)ys[j - 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (v[u] != 0.0) atomicAdd(s_acc + gm * kEcM + xy[u], v[u]);
      }
    }
    mb += (size_t)ncell * nM;
  }
  __syncthreads();
  double* out = out_ecnt + (size_t)e.pair * G * TAB_SZ;  // zero-filled by the caller; several workgroups per pair
  for (int q = threadIdx.x; q < G * kEcM; q += blockDim.x) {
    const double v = s_acc[q];
    if (v != 0.0) unsafeAtomicAdd(out + (size_t)(q / kEcM) * TAB_SZ + q % kEcM, v);
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
// Flag region (ints): forward ticket counters [kQueues], backward [kQueues],
// abort flag, padding to 4 x kQueues, then forward / backward / accumulator
// counters per band boundary.
constexpr int kFlagHead = 4 * kQueues;

// LDS of band_bwd_kernel<NT> before its dynamic emission-count tables.
size_t bwd_static_lds(int NT) {
  const size_t extra = 64;  // the ticket and slot words
  switch (NT) {
    case 2: return bwd_smem<2>() + extra;
    case 4: return bwd_smem<4>() + extra;
    case 6: return bwd_smem<6>() + extra;
    case 8: return bwd_smem<8>() + extra;
    default: return bwd_smem<10>() + extra;
  }
}

struct BPlan {
  bool ok = false;
  std::string why;
  int S = 0, NB = 0, NT = 0, NBP = 0, R = 0, e_idx = -1, G = 0;
  std::vector<int32_t> bidx, btype, bgrp;
  std::vector<int32_t> mrank, mgrp;     // [NBP] M rank of slot k; [NBP] group of M rank m
  std::vector<int32_t> qoff_f, qoff_b;  // kQueues + 1 each
  std::vector<BDesc> fwd, bwd;
  std::vector<EcPair> ec;               // binned emission counts: one entry per pair
  int nM = 0, ec_mode = 0;
  int nslots = 0;                       // checkpointed bands' lattice slots
  int64_t slot_stride = 0, slot0 = 0;   // doubles per slot; first slot (data region)
  int64_t ck_lat = 0, ck_saving = 0;    // checkpointed pairs' lattices; what the slots save of them
  int n_flags = 0;
  size_t idx_b = 0, desc_b = 0, pm_b = 0, tab_b = 0, pinfo_b = 0, flag_b = 0, data_b = 0;
  size_t total() const { return idx_b + desc_b + pm_b + tab_b + pinfo_b + flag_b + data_b; }
};

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Compute units per ticket queue (per XCD): the current device's count / 8,
// or 32 (MI355X: 256 / 8) where no device can be queried.
double cus_per_queue() {
  static int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      return n;
    (void)hipGetLastError();
    return 256;
  }();
  return std::max(1.0, cus / (double)kQueues);
}

// Pair groups per queue hold pairs until they can feed this many times the
// XCD's workgroup slots (TKFMI_BAND_FEED overrides the default for tuning).
double band_feed() {
  const char* e = getenv("TKFMI_BAND_FEED");
  const double v = e ? atof(e) : 0.0;
  return v > 0.0 ? v : 3.0;
}

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
  pl.R = 16 * band_rt(pl.NT);
  const int R = pl.R;
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
  const bool counts = !fwd_only && !(flags & TKFMI_NO_COUNTS);
  const int NBP = pl.NBP;
  // emission counts: binned (I and D types in LDS per band, M masses per
  // cell binned per pair) up to kEcMaxG groups, else atomics per cell
  const bool ecnt = (flags & TKFMI_WANT_EMIT_COUNTS) && !fwd_only && !dense;
  // mode 3 (every type binned in the backward's LDS) where the tables fit
  // beside the kernel's own LDS at its workgroups per CU; else mode 1 up to
  // kEcMaxG groups (the M reduction kernel's LDS), else atomics per cell
  const size_t lds_m3 = bwd_static_lds(pl.NT) + (size_t)pl.G * (A_PAD + pl.R + kEcM) * 8;
  const bool m3 = lds_m3 + 1024 <= (size_t)(160 * 1024) / band_occ(pl.NT);
  // mode 1's reduction kernel holds the tables and the longest pair's codes
  int max_xy = 0;
  for (int b = 0; b < B; ++b) max_xy = std::max(max_xy, p->h_x_len[b] + p->h_y_len[b]);
  const bool m1 = pl.G <= kEcMaxG && (size_t)pl.G * kEcM * 8 + (size_t)max_xy + 16 <= (size_t)(160 * 1024);
  pl.ec_mode = !ecnt ? 0 : (m3 ? 3 : (m1 ? 1 : 2));
  pl.mrank.assign(NBP, -1);
  pl.mgrp.assign(NBP, 0);
  for (int k = 0; k < pl.NB; ++k)
    if (pl.btype[k] == ST_M) { pl.mgrp[pl.nM] = pl.bgrp[k]; pl.mrank[k] = pl.nM++; }
  pl.idx_b = align256((5 * (size_t)NBP + 2 * (kQueues + 1)) * 4);
  pl.pm_b = align256((size_t)pm_stride(NBP) * 8 * p->n_trans);
  pl.tab_b = dense ? 0 : align256((size_t)p->n_tab * pl.G * TAB_SZ * 8);
  pl.pinfo_b = align256((size_t)std::max(B, 1) * 16);
  std::vector<int64_t> post_off(B);
  int64_t ap = 0;
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    if (lx < 0 || ly < 0 || lx + ly > 64000) { pl.why = "sequence lengths out of range"; return pl; }
    post_off[b] = ap;
    ap += (int64_t)(lx + 1) * (ly + 1) * p->S;
  }
  // per pair: lattice, band boundaries, count accumulator; flags
  int64_t off = 0;  // doubles into the data region
  int nbnd = 0;
  // A pair whose lattice exceeds the checkpoint budget (or every pair with
  // TKFMI_CHECKPOINT) keeps none: its backward bands recompute their own
  // lattices into slots (one per workgroup that can be resident), from the
  // forward's boundary rows, which a second set of boundary rows then keeps
  // intact for the backward's own pushes.
  // Slots: no more checkpointed bands are resident at once than there are
  // such bands, or workgroups the device holds (band_occ(NT) per CU).  When
  // the budget (not TKFMI_CHECKPOINT) selects pairs whose lattices the slots
  // would not halve (a few long pairs among short ones) and those lattices
  // total at most 16 budgets, they are kept whole.
  const size_t budget = tkfmi_checkpoint_budget_bytes();
  struct PairBands { int G; bool ck; int64_t lat, bnd0, bbnd0, acc, mb; int flag0; double cost; };
  std::vector<PairBands> pb(B);
  int64_t slot_stride = 0, ck_bands = 0, ck_lat = 0;
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    pb[b].G = (lx + R) / R;  // ceil((Lx + 1) / R)
    // band g's lattice: T_g = Ly + nr_g steps of R rows
    const int64_t lat_d = ((int64_t)pb[b].G * ly + lx + 1) * R * (NBP + 1);
    pb[b].ck = !fwd_only && ((flags & TKFMI_CHECKPOINT) || (budget > 0 && (size_t)lat_d * 8 > budget));
    if (pb[b].ck) {
      slot_stride = std::max<int64_t>(slot_stride, ((int64_t)ly + std::min(R, lx + 1)) * R * (NBP + 1));
      ck_bands += pb[b].G;
      ck_lat += lat_d;
    }
  }
  int nslots = (int)std::min<int64_t>(ck_bands, (int64_t)(cus_per_queue() * kQueues) * band_occ(pl.NT));
  if (!(flags & TKFMI_CHECKPOINT) && ck_lat > 0 && 2 * (ck_lat - (int64_t)nslots * slot_stride) < ck_lat &&
      (size_t)ck_lat * 8 <= 16 * budget) {
    for (int b = 0; b < B; ++b) pb[b].ck = false;
    nslots = 0;
    slot_stride = ck_lat = 0;
  }
  for (int b = 0; b < B; ++b) {
    const int lx = p->h_x_len[b], ly = p->h_y_len[b];
    const int G = pb[b].G;
    const int64_t lat_d = ((int64_t)G * ly + lx + 1) * R * (NBP + 1);
    pb[b].lat = (fwd_only || pb[b].ck) ? -1 : off;
    if (pb[b].lat >= 0) off += lat_d;
    pb[b].bnd0 = off;
    off += (int64_t)(G - 1) * (ly + 1) * (NBP + 1);
    pb[b].bbnd0 = pb[b].bnd0;
    if (pb[b].ck) {
      pb[b].bbnd0 = off;
      off += (int64_t)(G - 1) * (ly + 1) * (NBP + 1);
    }
    pb[b].acc = off;
    if (counts && G > 1) off += (int64_t)NBP * NBP;
    pb[b].mb = off;
    if (pl.ec_mode == 1) off += ((int64_t)G * ly + lx + 1) * R * pl.nM;
    pb[b].flag0 = kFlagHead + 3 * nbnd;
    nbnd += G - 1;
    pb[b].cost = (double)G * (ly + R);
  }
  pl.nslots = nslots;
  pl.ck_saving = ck_lat - (int64_t)pl.nslots * slot_stride;
  pl.ck_lat = ck_lat;
  pl.slot_stride = slot_stride;
  pl.slot0 = off;
  off += (int64_t)pl.nslots * slot_stride;
  pl.n_flags = kFlagHead + 3 * nbnd + pl.nslots;
  pl.flag_b = align256((size_t)pl.n_flags * 4);
  // Pairs to queues (one per XCD), longest first to the least loaded queue.
  // Within a queue, the pairs are taken in groups just large enough to keep
  // the XCD's compute units busy: a band can start about lag = R + CQ steps
  // after the band it reads from, so one pair can feed T / lag compute units
  // (T: steps per band), and a group holds pairs until the sum reaches 3
  // times the XCD's workgroup slots (swept on 1024-pair batches at 32, 52,
  // 82 and 162 states: 3 was fastest or within 0.5 % of it everywhere;
  // 1.5 cost 1-3 %, 20 up to 13 %).  A group's bands are listed layer by
  // layer (band 0 of each pair, then band 1, ...; the backward from the
  // last band), so that a band's ticket comes about when its input is
  // ready, and a few pairs at a time share the XCD's L2.
  std::vector<int> order(B);
  for (int b = 0; b < B; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return pb[a].cost > pb[c].cost; });
  std::vector<std::vector<int>> queue(kQueues);
  std::vector<double> load(kQueues, 0.0);
  for (int b : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    queue[x].push_back(b);
    load[x] += pb[b].cost;
  }
  const double lag = R + CQ, feed = band_feed() * cus_per_queue() * band_occ(pl.NT);
  pl.qoff_f.assign(kQueues + 1, 0);
  pl.qoff_b.assign(kQueues + 1, 0);
  std::vector<BDesc> fwd_all, bwd_all;  // per pair, bands 0..G-1
  std::vector<int64_t> first(B, 0);
  for (int b = 0; b < B; ++b) {
    {
      const int lx = p->h_x_len[b], ly = p->h_y_len[b], G = pb[b].G;
      const int64_t bsz = (int64_t)(ly + 1) * (NBP + 1);
      const int fl = pb[b].flag0;
      first[b] = (int64_t)fwd_all.size();
      int64_t lat_band = pb[b].lat, mb_band = pb[b].mb;
      if (pl.ec_mode == 1) {  // the pair's bands in chunks of about kEcChunk masses
        EcPair e;
        memset(&e, 0, sizeof(e));
        e.pair = b; e.Lx = lx; e.Ly = ly; e.G = G;
        e.x_off = p->h_x_off[b]; e.y_off = p->h_y_off[b];
        int64_t mo = pb[b].mb, vol = 0;
        e.g0 = 0;
        e.mb_off = mo;
        for (int g = 0; g < G; ++g) {
          const int64_t band_vol = (int64_t)(ly + std::min(R, lx + 1 - g * R)) * R * pl.nM;
          vol += band_vol;
          mo += band_vol;
          if (vol >= kEcChunk || g == G - 1) {
            e.g1 = g + 1;
            pl.ec.push_back(e);
            e.g0 = g + 1;
            e.mb_off = mo;
            vol = 0;
          }
        }
      }
      for (int g = 0; g < G; ++g) {
        BDesc d;
        memset(&d, 0, sizeof(d));
        d.pair = b; d.Lx = lx; d.Ly = ly; d.i0 = g * R; d.nr = std::min(R, lx + 1 - g * R);
        d.g = g; d.G = G; d.T = ly + d.nr;
        d.lat_off = lat_band;
        if (lat_band >= 0) lat_band += (int64_t)d.T * R * (NBP + 1);
        d.mb_off = mb_band;
        mb_band += (int64_t)d.T * R * pl.nM;
        d.pm_off = (p->n_trans > 1 ? (int64_t)b : 0) * pm_stride(NBP);
        d.tab_off = (!dense && p->n_tab > 1 ? (int64_t)b : 0) * pl.G * TAB_SZ;
        d.x_off = p->h_x_off[b]; d.y_off = p->h_y_off[b];
        d.post_off = post_off[b];
        d.acc_off = pb[b].acc;
        // forward: boundary g-1 in, g out
        d.bin_off = g > 0 ? pb[b].bnd0 + (int64_t)(g - 1) * bsz : -1;
        d.fin = g > 0 ? fl + 3 * (g - 1) : -1;
        d.bout_off = g < G - 1 ? pb[b].bnd0 + (int64_t)g * bsz : -1;
        d.fout = g < G - 1 ? fl + 3 * g : -1;
        d.ain = d.aout = -1;
        d.fbin_off = -1;
        fwd_all.push_back(d);
        // backward: boundary g in, g-1 out (over the forward's boundary rows
        // once the forward launch has ended, unless they are kept for the
        // recomputation of a checkpointed pair)
        d.fbin_off = (pb[b].ck && g > 0) ? pb[b].bnd0 + (int64_t)(g - 1) * bsz : -1;
        d.bin_off = g < G - 1 ? pb[b].bbnd0 + (int64_t)g * bsz : -1;
        d.fin = g < G - 1 ? fl + 3 * g + 1 : -1;
        d.bout_off = g > 0 ? pb[b].bbnd0 + (int64_t)(g - 1) * bsz : -1;
        d.fout = g > 0 ? fl + 3 * (g - 1) + 1 : -1;
        d.ain = (counts && g < G - 1) ? fl + 3 * g + 2 : -1;
        d.aout = (counts && g > 0) ? fl + 3 * (g - 1) + 2 : -1;
        bwd_all.push_back(d);
      }
    }
  }
  for (int x = 0; x < kQueues; ++x) {
    const std::vector<int>& qp = queue[x];
    size_t a = 0;
    while (a < qp.size()) {
      size_t e = a;
      double fed = 0.0;
      int maxG = 0;
      while (e < qp.size() && (e == a || fed < feed)) {
        fed += (p->h_y_len[qp[e]] + R) / lag;
        maxG = std::max(maxG, pb[qp[e]].G);
        ++e;
      }
      for (int layer = 0; layer < maxG; ++layer)
        for (size_t q = a; q < e; ++q) {
          const int b = qp[q], G = pb[b].G;
          if (layer >= G) continue;
          pl.fwd.push_back(fwd_all[first[b] + layer]);
          pl.bwd.push_back(bwd_all[first[b] + G - 1 - layer]);
        }
      a = e;
    }
    pl.qoff_f[x + 1] = (int32_t)pl.fwd.size();
    pl.qoff_b[x + 1] = (int32_t)pl.bwd.size();
  }
  pl.desc_b = align256(sizeof(BDesc) * std::max<size_t>(pl.fwd.size() + pl.bwd.size(), 1) +
                       sizeof(EcPair) * pl.ec.size());
  pl.data_b = (size_t)off * 8;
  // offsets of the data region, relative to ws (in doubles)
  const int64_t data0 = (int64_t)(pl.idx_b + pl.desc_b + pl.pm_b + pl.tab_b + pl.pinfo_b + pl.flag_b) / 8;
  const int64_t pm0 = (int64_t)(pl.idx_b + pl.desc_b) / 8;
  const int64_t tab0 = (int64_t)(pl.idx_b + pl.desc_b + pl.pm_b) / 8;
  for (std::vector<BDesc>* v : {&pl.fwd, &pl.bwd})
    for (BDesc& d : *v) {
      if (d.lat_off >= 0) d.lat_off += data0;
      if (d.fbin_off >= 0) d.fbin_off += data0;
      if (d.bin_off >= 0) d.bin_off += data0;
      if (d.bout_off >= 0) d.bout_off += data0;
      d.acc_off += data0;
      d.mb_off += data0;
      d.pm_off += pm0;
      d.tab_off += tab0;
    }
  for (EcPair& e : pl.ec) e.mb_off += data0;
  pl.slot0 += data0;
  pl.ok = true;
  return pl;
}

template <int NT, bool DENSE, bool CK>
int launch_bwd(const BParams& prm, int n, size_t dyn, hipStream_t s, std::string& err) {
  if (dyn > 0) {
    const hipError_t ea = hipFuncSetAttribute((const void*)band_bwd_kernel<NT, DENSE, CK>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (ea != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(ea); return TKFMI_EHIP; }
  }
  hipLaunchKernelGGL((band_bwd_kernel<NT, DENSE, CK>), dim3(n), dim3(kThr), dyn, s, prm);
  return TKFMI_OK;
}

template <int NT, bool DENSE>
int launch_nt(bool bwd, const BParams& prm, int n, size_t dyn, hipStream_t s, std::string& err) {
  if (n == 0) return TKFMI_OK;
  if (bwd) {
    const int rc = prm.nslots > 0 ? launch_bwd<NT, DENSE, true>(prm, n, dyn, s, err)
                                  : launch_bwd<NT, DENSE, false>(prm, n, dyn, s, err);
    if (rc != TKFMI_OK) return rc;
  } else {
    hipLaunchKernelGGL((band_fwd_kernel<NT, DENSE>), dim3(n), dim3(kThr), 0, s, prm);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("band launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

template <bool DENSE>
int launch_d(int NT, bool bwd, const BParams& prm, int n, size_t dyn, hipStream_t s, std::string& err) {
  switch (NT) {
    case 2: return launch_nt<2, DENSE>(bwd, prm, n, dyn, s, err);
    case 4: return launch_nt<4, DENSE>(bwd, prm, n, dyn, s, err);
    case 6: return launch_nt<6, DENSE>(bwd, prm, n, dyn, s, err);
    case 8: return launch_nt<8, DENSE>(bwd, prm, n, dyn, s, err);
    default: return launch_nt<10, DENSE>(bwd, prm, n, dyn, s, err);
  }
}

}  // namespace tkfmi_band

using namespace tkfmi_band;

bool tkfmi_band_applies(const tkfmi_problem* p, uint32_t flags) {
  if (flags & TKFMI_CLASS_KERNEL) return false;  // the one-workgroup-per-pair kernel
  const BPlan pl = plan(p, flags);
  // checkpointed pairs: the slots (a band lattice per resident workgroup)
  // must at least halve their lattices; otherwise (a few very long pairs)
  // the one-workgroup-per-pair kernel's K-antidiagonal checkpoints are used
  return pl.ok && (pl.ck_lat == 0 || 2 * pl.ck_saving >= pl.ck_lat);
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
  memcpy(blob.data() + 3 * pl.NBP * 4, pl.qoff_f.data(), (kQueues + 1) * 4);
  memcpy(blob.data() + (3 * pl.NBP + kQueues + 1) * 4, pl.qoff_b.data(), (kQueues + 1) * 4);
  memcpy(blob.data() + (3 * pl.NBP + 2 * (kQueues + 1)) * 4, pl.mrank.data(), pl.NBP * 4);
  memcpy(blob.data() + (4 * pl.NBP + 2 * (kQueues + 1)) * 4, pl.mgrp.data(), pl.NBP * 4);
  memcpy(blob.data() + pl.idx_b, pl.fwd.data(), sizeof(BDesc) * nf);
  memcpy(blob.data() + pl.idx_b + sizeof(BDesc) * nf, pl.bwd.data(), sizeof(BDesc) * nbw);
  memcpy(blob.data() + pl.idx_b + sizeof(BDesc) * (nf + nbw), pl.ec.data(), sizeof(EcPair) * pl.ec.size());
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
  prm.e_idx = pl.e_idx; prm.G = pl.G; prm.nM = pl.nM; prm.ec_mode = pl.ec_mode;
  prm.mrank = d_idx + 3 * pl.NBP + 2 * (kQueues + 1);
  {
    const char* e = getenv("TKFMI_CHAIN_SPIN_LIMIT");
    prm.spin_max = e ? std::max(0, atoi(e)) : kSpinMax;
  }
  prm.bidx = d_idx; prm.btype = d_idx + pl.NBP; prm.bgrp = d_idx + 2 * pl.NBP;
  prm.x_cat = p->x_cat; prm.y_cat = p->y_cat; prm.log_emit = p->log_emit;
  prm.ws = wsd;
  prm.flagv = flagv; prm.abortf = flagv + 2 * kQueues;
  prm.nslots = pl.nslots;
  prm.slots = wsd + pl.slot0;
  prm.slot_stride = pl.slot_stride;
  prm.slotf = flagv + (pl.n_flags - pl.nslots);
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
    pc.ticket = flagv + dir * kQueues;
    pc.qoff = d_idx + 3 * pl.NBP + dir * (kQueues + 1);
    pc.desc = reinterpret_cast<const BDesc*>(wsb + pl.idx_b) + (dir ? nf : 0);
    const std::string name = std::string(dir ? "band_bwd_kernel<" : "band_fwd_kernel<") + std::to_string(pl.NT) +
                             (dir && pl.nslots > 0 ? ",recompute>" : ">");
    const int slot = tkfmi_ktime_start(name.c_str(), B, cells, stream, err);
    if (slot == -2) return TKFMI_EHIP;
    const size_t dyn = (dir == 1 && pl.ec_mode == 1)   ? (size_t)pl.G * (A_PAD + pl.R) * 8
                       : (dir == 1 && pl.ec_mode == 3) ? (size_t)pl.G * (A_PAD + pl.R + kEcM) * 8
                                                       : 0;
    rc = dense ? launch_d<true>(pl.NT, dir == 1, pc, (int)(dir ? nbw : nf), dyn, stream, err)
               : launch_d<false>(pl.NT, dir == 1, pc, (int)(dir ? nbw : nf), dyn, stream, err);
    if (rc != TKFMI_OK) return rc;
    if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return TKFMI_EHIP;
  }
  if (pl.ec_mode == 1 && !pl.ec.empty()) {  // M-type emission counts, binned per pair
    int max_xy = 0;
    for (const EcPair& q : pl.ec) max_xy = std::max(max_xy, q.Lx + q.Ly);
    const size_t smem = (size_t)pl.G * kEcM * 8 + (((size_t)max_xy + 15) & ~(size_t)15);
    hipError_t e = hipFuncSetAttribute((const void*)band_ecnt_m_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)smem);
    if (e != hipSuccess) { err = std::string("hipFuncSetAttribute: ") + hipGetErrorString(e); return TKFMI_EHIP; }
    const int slot = tkfmi_ktime_start("band_ecnt_m_kernel", B, cells, stream, err);
    if (slot == -2) return TKFMI_EHIP;
    hipLaunchKernelGGL(band_ecnt_m_kernel, dim3((unsigned)pl.ec.size()), dim3(512), smem, stream,
                       reinterpret_cast<const EcPair*>(wsb + pl.idx_b + sizeof(BDesc) * (nf + nbw)), pl.R, pl.nM,
                       pl.G, d_idx + 4 * pl.NBP + 2 * (kQueues + 1), p->x_cat, p->y_cat, wsd, pinfo, o->emit_counts);
    if (hipGetLastError() != hipSuccess) { err = "emission-count launch failed"; return TKFMI_EHIP; }
    if (tkfmi_ktime_stop(slot, stream, err) != TKFMI_OK) return TKFMI_EHIP;
  }
  if (o->status) {
    hipLaunchKernelGGL(raise_status_kernel, dim3(1), dim3(64), 0, stream, prm.abortf, o->status);
    if (hipGetLastError() != hipSuccess) { err = "status launch failed"; return TKFMI_EHIP; }
  }
  return TKFMI_OK;
}
