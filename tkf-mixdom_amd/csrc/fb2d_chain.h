// tkfmi: device pieces shared by the chained cell-per-lane kernels
// (fb2d_chain.hip: forward and backward as separate launches;
// fb2d_bidi.hip: forward and backward of a pair in one launch, meeting at
// an antidiagonal): descriptors, bounded waits, the boundary-row ring,
// tickets, the LDS layout, input staging and the count wave.
#ifndef TKFMI_FB2D_CHAIN_H
#define TKFMI_FB2D_CHAIN_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fb2d_cell.h"
#include "tkfmi.h"

namespace tkfmi_v2 {

constexpr int CQ = 16;                 // boundary entries per published / fetched chunk
// Backward count staging per recursion wave: its 64 cells' F-scaled and g
// vectors, 16 doubles each (16 KB).
constexpr int kStageDoubles = 2048;
constexpr int kSpinMax = 1 << 22;      // bounded waits (each spin sleeps ~0.2 us + a load)
// TKFMI_CHAIN_SPIN_LIMIT overrides kSpinMax; tests set it tiny so that
// waits time out and the abort path runs with every kernel still ending.

struct CDesc {
  int32_t pair, Lx, Ly, band0;  // pair geometry; first band of this workgroup
  int32_t nw, g, G, T;          // waves holding a band (<= W), chain index, chain length, steps
  int64_t x_off, y_off;
  int64_t lat_off, post_off, pm_off;
  int64_t xin_off, xout_off;    // doubles into the boundary region (-1: none)
  int32_t fin, fout;            // progress counters (ints into the flag region; -1: none)
  // Forward and backward in one launch (fb2d_bidi.hip); unused otherwise.
  int32_t dir;                  // 0: forward workgroup, 1: backward workgroup
  int32_t C;                    // meeting antidiagonal of the pair (i + j)
  int32_t cnt, bflag;           // flag indices: the pair's partial counter, this band group's cut flag
  int32_t pslot, gpair;         // the pair's first partial slot (doubles), forward groups of the pair
  int64_t cut_off;              // doubles into the cut buffer: [2][16][Lx + 1]
};

struct CParams {
  int32_t S, flags, n_tab, G_tab;
  int32_t gmap[16];  // layout group -> caller table group
  const double* log_tab;
  const uint8_t* x_cat;
  const uint8_t* y_cat;
  double* lattice;
  double* xring;      // boundary rows between the workgroups of a chain
  int32_t* flagv;     // progress counters
  int32_t* ticket;    // this launch's ticket counter
  int32_t* abortf;    // raised by a timed-out wait
  int32_t spin_max;   // spins before a wait times out
  double* pinfo;      // [B][2]: mantissa and binary exponent of P(x, y)
  double* out_logP;
  double* out_n;
  double* out_post;
  double* out_ecnt;
  double* cutbuf;     // backward values at antidiagonals C - 1 and C, per pair
  double* pslot;      // per forward group of a pair: mantissa, exponent of its partial of P
  int32_t fail_base;  // one-launch schedule: flagv[fail_base + pair] = 1 when a wait of the pair timed out
};

// Waits (whole wave, uniform) until *flag >= need, then acquires (agent
// scope) what was released with the flag.  Returns false when the launch was
// aborted, by this wait timing out or by another one.  The flag is polled
// with relaxed loads and the acquire is one fence after the last: an
// agent-scope acquire invalidates the XCD's L2, and a poll that acquired
// would invalidate it for every workgroup on the XCD once per spin.
__device__ inline bool wait_ge(int* flag, int need, int* abortf, int spin_max) {
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
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Before one thread releases, at agent scope, data that every wave of the
// workgroup wrote: each wave's own release fence (its stores complete, the
// XCD's L2 written back), then the barrier.  A barrier alone does not wait
// for the other waves' stores.
__device__ __forceinline__ void wg_release_agent() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
}

// A cell-step's lattice slot: its NB state values and binary exponent as
// NB + 1 doubles in pairs, in the cell-step's region of SLOTS x 64 doubles:
// the 64 lanes' pair q (elements 2q, 2q + 1) are 64 contiguous 16-byte
// entries, so each pair is one 1 KB store or load of the wave (8-byte
// elements took twice the memory instructions: on the 1024-pair K=3 batch
// the one-launch longest pairs went 5.27 -> 5.10 ms and the main backward
// launch 4.61 -> 4.49 ms).  `p` points at the lane's pair 0
// (the region + 2 lane doubles); element b is at p + 128 (b / 2) + b % 2,
// the exponent at element NB.
// NT: non-temporal, for the two-launch schedule, where the lattice is
// written once by the forward launch and read once by the backward launch:
// held in the XCD's L2 it only evicted what is reused there (the boundary
// rows and the other launches' working sets); streamed, that batch's main
// forward launch went from 3.36 to 2.99 ms and the call from 8.1 to 7.7 ms.
// The one-launch schedule reads back what it wrote within microseconds,
// from L2: ordinary stores there (streamed, c1's longest pair alone took
// 4.3 instead of 4.0 ms).
typedef double slot_d2 __attribute__((ext_vector_type(2)));

template <int NB, bool NT>
__device__ __forceinline__ void store_slot(double* p, const double* v, int E) {
  constexpr int NP = (NB + 2) / 2;  // pairs of the NB + 1 elements
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int b0 = 2 * q, b1 = 2 * q + 1;
    slot_d2 x;
    x.x = b0 == NB ? (double)E : v[b0 < NB ? b0 : 0];
    x.y = b1 == NB ? (double)E : (b1 < NB ? v[b1] : 0.0);
    slot_d2* d = reinterpret_cast<slot_d2*>(p + 128 * q);
    if constexpr (NT) __builtin_nontemporal_store(x, d);
    else *d = x;
  }
}

// Reads one cell's lattice slot (store_slot's layout) into v[0..NB) and
// the exponent into v[SLOTS - 1].
template <int NB, bool NT>
__device__ __forceinline__ void load_slot(const double* p, double* v) {
  constexpr int NP = (NB + 2) / 2;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const slot_d2* s = reinterpret_cast<const slot_d2*>(p + 128 * q);
    slot_d2 x;
    if constexpr (NT) x = __builtin_nontemporal_load(s);
    else x = *s;
    const int b0 = 2 * q, b1 = 2 * q + 1;
    if (b0 < NB) v[b0] = x.x; else v[SLOTS - 1] = x.x;
    if (b1 < NB) v[b1] = x.y; else if (b1 == NB) v[SLOTS - 1] = x.y;
  }
}

// Boundary entries between the workgroups of a chain are written and read
// as device-scope relaxed atomics, which gfx950 performs at the device's
// coherence point (past the per-XCD L2s).  Publishing then only has to wait
// until this lane's entry stores have completed before it raises the counter:
// an agent-scope release fence would also write back every dirty line of the
// XCD's L2, and there the forward's lattice stores keep megabytes dirty, so
// each publish would wait for a flush of the whole cache.
__device__ __forceinline__ void ring_store(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes `count` boundary entries (this lane's ring_store calls before it)
// to the next workgroup of the chain.
__device__ __forceinline__ void publish(int* flag, int count) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the entry stores have completed
  __hip_atomic_store(flag, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Copies boundary entries [e, e_end) into the LDS chunk buffer (whole wave),
// after wait_ge has acquired their progress counter.
__device__ __forceinline__ void fetch_chunk(double* s_xc, const double* src, int n) {
  for (int q = threadIdx.x & 63; q < n; q += 64)
    s_xc[q] = __hip_atomic_load(src + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  wave_lds_sync();
}

// Writes one boundary entry (the M and D parts of v, then the exponent) to
// the LDS ring of the next wave or, for the next workgroup, to global memory.
template <class L, int NX>
__device__ __forceinline__ void put_entry(double* e, bool global, const double* v, int E) {
  constexpr int NB = L::NB, nM = count_type<L, ST_M>();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int o = -1;
    if (L::type(b) == ST_M) o = rank_in_type<L, ST_M>(b);
    if (L::type(b) == ST_D) o = nM + rank_in_type<L, ST_D>(b);
    if (o >= 0) {
      if (global) ring_store(e + o, v[b]);
      else e[o] = v[b];
    }
  }
  if (global) ring_store(e + NX - 1, (double)E);
  else e[NX - 1] = (double)E;
}

// The launch's next ticket, the same for the whole workgroup.  A ticket past
// the grid means the counter was not zeroed before the launch (a caller
// error, e.g. a workspace shared by overlapping calls): the workgroup raises
// the abort flag and its waves end here (every wave sees the same ticket),
// without touching a descriptor.
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

// LDS layout (doubles): tables | start / end transitions | emission counts
// (backward) | short rings | chunk buffer | count staging (backward, 64 cells
// x 32 doubles per recursion wave) | bytes: x rows of the workgroup, y.
template <class L, int W>
struct CLds {
  static constexpr int NX = count_type<L, ST_M>() + count_type<L, ST_D>() + 1;
  static constexpr int tab = 0;
  static constexpr int misc = L::NG * TAB_SZ;  // p_start[NB], p_end[NB], P[S->E]
  static constexpr int ecnt = misc + 32;
  static constexpr int ring(bool bwd) { return ecnt + (bwd ? L::NG * TAB_SZ : 0); }
  static constexpr int xc(bool bwd) { return ring(bwd) + (W - 1) * 2 * NX; }
  static constexpr int stage(bool bwd) { return ((xc(bwd) + CQ * NX) + 1) & ~1; }
  static constexpr int seq_bytes(bool bwd) { return 8 * (stage(bwd) + (bwd ? W * kStageDoubles : 0)); }
  static constexpr int xbytes = 64 * W + 8;
  static size_t total(bool bwd, int max_ly) { return (size_t)seq_bytes(bwd) + xbytes + ((max_ly + 8) & ~7) + 16; }
};

// Stages the exponentiated emission tables and the workgroup's x rows / y.
// The start / end transitions go to LDS too: a global load inside the sweep
// would make the compiler wait for every outstanding lattice store (stores
// and loads share one counter) at the next use of its register.
template <class L, int W, int NTHR = W * 64>
__device__ __forceinline__ void stage_inputs(const CParams& Pr, const CDesc& pd, const double* pm,
                                             double* s_tab, uint8_t* s_x, uint8_t* s_y) {
  constexpr int NG = L::NG, NB = L::NB;
  static_assert(2 * NB + 1 <= 32, "start / end staging");
  double* s_misc = s_tab + NG * TAB_SZ;
  for (int q = threadIdx.x; q < 2 * NB + 1; q += NTHR) s_misc[q] = pm[2 * NB * NB + q];
  const double* ltab = Pr.log_tab + (Pr.n_tab > 1 ? (size_t)pd.pair * Pr.G_tab * TAB_SZ : 0);
  for (int q = threadIdx.x; q < NG * TAB_SZ; q += NTHR) {
    const int g = q / TAB_SZ, o = q - g * TAB_SZ;
    s_tab[q] = exp(ltab[Pr.gmap[g] * TAB_SZ + o]);
  }
  // s_x[q] = x[64 band0 + q - 1] (the character of row 64 band0 + q)
  const int xb = 64 * pd.band0 - 1;
  for (int q = threadIdx.x; q < 64 * W + 1; q += NTHR) {
    const int xi = xb + q;
    s_x[q] = (xi >= 0 && xi < pd.Lx) ? (uint8_t)min((int)Pr.x_cat[pd.x_off + xi], A_PAD - 1) : 0;
  }
  for (int q = threadIdx.x; q < pd.Ly; q += NTHR) s_y[q] = min((int)Pr.y_cat[pd.y_off + q], A_PAD - 1);
  if (threadIdx.x == 0) s_y[pd.Ly] = 0;
}



// Count wave, one step: the expected-count MFMAs of the cells the recursion
// waves staged at the previous step.  On gfx950 an FP64 MFMA occupies its
// SIMD's FP64 pipe for 64 cycles, with no overlap with VALU FP64 work, so a
// separate wave (on its own SIMD) keeps them off the recursion's dependent
// chain.  Each stage (64 cells) is read in two halves of 32 cells into
// registers; the reads of the next half are issued before the 8 MFMAs of the
// current one (4 cells each, two accumulator chains: acc (+acc2) += F^T g),
// so LDS latency hides behind the matrix pipe.  When both halves of stage v
// are in registers the wave releases it (s_cflag[v] = t): recursion wave v
// may then overwrite it with its cells of step t.  Stages of waves that were
// not active at step t - 1, or that staged nothing there (steps before t_lo),
// are read but not counted.
template <int W>
__device__ __forceinline__ void count_step(const double* s_stage, int NW, int Ly, int t, int lane, int* s_cflag,
                                           d4& acc, d4& acc2, int t_lo = 0) {
  int ln = lane;
  asm volatile("" : "+v"(ln));  // recompute the addresses here (hoisted, they would take 48 VGPRs)
  auto load = [&](int i, double (&a)[8], double (&b)[8]) {
    const double* sf = s_stage + (size_t)(i >> 1) * kStageDoubles;
    const double* sg = sf + kStageDoubles / 2;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = 4 * (8 * (i & 1) + k) + (ln >> 4);
      const int o = c * 16 + ((ln & 15) ^ (c & 15));
      a[k] = sf[o];
      b[k] = sg[o];
    }
  };
  double ca[8], cb[8], na[8], nb[8];
  load(0, ca, cb);
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    const int v = i >> 1;
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // half i is in registers
    if ((i & 1) && lane == 0) __hip_atomic_store(s_cflag + v, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (i + 1 < 2 * W) load(i + 1, na, nb);
    __builtin_amdgcn_sched_barrier(0);
    if (v < NW && t >= 64 * v + 1 && t <= 64 * v + 64 + Ly && t - 1 >= t_lo) {  // wave v staged at step t - 1
#pragma unroll
      for (int k = 0; k < 8; k += 2) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[k], cb[k], acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[k + 1], cb[k + 1], acc2, 0, 0, 0);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) { ca[k] = na[k]; cb[k] = nb[k]; }
  }
}

// Launches the one-launch forward + backward kernel (fb2d_bidi.hip) of
// layout pl with W recursion waves per workgroup over n descriptors.
int launch_bidi(const Plan& pl, int W, const CParams& prm, const double* pmat, const CDesc* d, int n,
                int max_ly, int wpc, hipStream_t s, std::string& err);
// Sets out_logP[b] = NaN for every pair b with fail[b] != 0.
int launch_mark_failed(const int32_t* fail, int B, double* out_logP, hipStream_t s, std::string& err);

}  // namespace tkfmi_v2

#endif
