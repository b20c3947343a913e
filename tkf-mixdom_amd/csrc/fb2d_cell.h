// tkfmi: pieces shared by the cell-per-lane forward/backward kernels
// (one workgroup per pair in fb2d_v2.hip, chained workgroups per pair in
// fb2d_chain.hip): state layouts, the probability-space transition block,
// DPP / LDS / scalar-cache helpers and the host-side layout detection.
#ifndef TKFMI_FB2D_CELL_H
#define TKFMI_FB2D_CELL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "tkfmi.h"

namespace tkfmi_v2 {

constexpr int ST_S = 0, ST_M = 1, ST_I = 2, ST_D = 3, ST_E = 4;
constexpr int A_PAD = TKFMI_A_PAD;
constexpr int TAB_SZ = TKFMI_TAB_SZ;
constexpr int EXP_NONE = -(1 << 29);
constexpr int SLOTS = 16;  // doubles stored per cell: NB mantissas, exponent in slot 15
constexpr double LN2 = 0.69314718055994530942;

typedef double d4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ state layouts
struct LTkf {  // S, M, I, D, E
  static constexpr int NB = 3, NG = 1, SOFF = 1, EIDX = 4;
  static constexpr int type(int b) { return b == 0 ? ST_M : (b == 1 ? ST_I : ST_D); }
  static constexpr int group(int) { return 0; }
};
template <int K, int F>
struct LMix {  // SS, EE, body 2 + d*5F + uv*F + f, uv in (MM, MI, MD, II, DD)
  static constexpr int NB = 5 * K * F, NG = K * F, SOFF = 2, EIDX = 1;
  static constexpr int type(int b) {
    const int uv = (b % (5 * F)) / F;
    return uv == 0 ? ST_M : ((uv == 1 || uv == 3) ? ST_I : ST_D);
  }
  static constexpr int group(int b) { return (b / (5 * F)) * F + (b % F); }
};

template <class L, int T>
constexpr int count_type() {
  int n = 0;
  for (int b = 0; b < L::NB; ++b) n += (L::type(b) == T);
  return n;
}
template <class L, int T>
constexpr int rank_in_type(int b) {  // index of body state b among states of type T
  int n = 0;
  for (int q = 0; q < b; ++q) n += (L::type(q) == T);
  return n;
}


// Per pair transition block (probability space), doubles:
//   [0, NB^2)          P[s][k]   (forward: row s contiguous)
//   [NB^2, 2NB^2)      PT[k'][k] = P[k][k'] (backward: row k' contiguous)
//   [2NB^2, +NB)       p_start[k] = P[S -> k]
//   [.., +NB)          p_end[k]   = P[k -> E]
//   [.., +1)           P[S -> E]
template <int NB>
constexpr int pm_stride() {
  return ((2 * NB * NB + 2 * NB + 1) + 15) / 16 * 16;
}

template <class L>
__global__ void prep_kernel(const double* __restrict__ log_trans, int S, int n_trans,
                            double* __restrict__ pmat) {
  constexpr int NB = L::NB;
  const int b = blockIdx.x;
  const double* lt = log_trans + (size_t)b * S * S;
  double* pm = pmat + (size_t)b * pm_stride<NB>();
  for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
    const int s = q / NB, k = q % NB;
    const double v = exp(lt[(s + L::SOFF) * S + (k + L::SOFF)]);
    pm[s * NB + k] = v;
    pm[NB * NB + k * NB + s] = v;
  }
  for (int k = threadIdx.x; k < NB; k += blockDim.x) {
    pm[2 * NB * NB + k] = exp(lt[k + L::SOFF]);
    pm[2 * NB * NB + NB + k] = exp(lt[(k + L::SOFF) * S + L::EIDX]);
  }
  if (threadIdx.x == 0) pm[2 * NB * NB + 2 * NB] = exp(lt[L::EIDX]);
  (void)n_trans;
}

// --------------------------------------------------------------- helpers
// DPP lane shift; a lane whose source is outside the wave reads 0
// (bound_ctrl), so no zero-initialised destination is needed.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = dpp_i32<CTRL>(__double2loint(v));
  const int hi = dpp_i32<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane l <- lane l-1
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane l <- lane l+1

__device__ __forceinline__ int frexp_exp(double v) { return __builtin_amdgcn_frexp_exp(v); }

// acc += bcast_K(prow) * x, where bcast_K reads lane K of the lane's row of
// 16 (DPP row_newbcast): with prow[lane] = P[s][lane % 16], every lane gets
// P[s][K] without a memory access.  One v_fmac_f64 with a 64-bit DPP source;
// prow must not be written by a VALU instruction right before (it is a
// loop-invariant register here).
template <int K>
__device__ __forceinline__ void fmac_bcast(double& acc, double prow, double x) {
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(prow), "v"(x), "i"(K));
}

// One row s of acc[k] += bcast_k(prow[s]) * x[s], k < NB (consecutive FMAs
// update different accumulators).
template <int NB, int K = 0>
__device__ __forceinline__ void row_fmac(double (&acc)[NB], double prow_s, double x_s) {
  if constexpr (K < NB) {
    fmac_bcast<K>(acc[K], prow_s, x_s);
    row_fmac<NB, K + 1>(acc, prow_s, x_s);
  }
}

// acc[k] += sum_s bcast_k(prow[s]) * x[s] for k, s < NB.
template <int NB>
__device__ __forceinline__ void row_contract(double (&acc)[NB], const double (&prow)[NB], const double (&x)[NB]) {
#pragma unroll
  for (int s = 0; s < NB; ++s) row_fmac<NB>(acc, prow[s], x[s]);
}

// Returns v through a VALU move: the result is not tied to the load that
// produced v, so using it inside a loop does not make the compiler wait for
// the memory counter there.
__device__ __forceinline__ double valu_copy(double v) {
  double r;
  asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// Hides a (uniform) pointer's value from the optimiser so that loads through
// it are not hoisted out of the step loop: the NB^2 transition values would
// not fit in SGPRs and would be spilled.  Each row is instead re-read through
// the scalar cache right before its FMAs.
typedef const __attribute__((address_space(4))) double* const_dptr;
__device__ __forceinline__ const_dptr launder(const double* p) {
  asm volatile("" : "+s"(p));
  return (const_dptr)p;  // constant address space: read with s_load
}

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Orders this wave's LDS writes before its later LDS reads of other lanes'
// data (the LDS executes a wave's operations in order; this stops the
// compiler from reordering them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ int wrap_mod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// ------------------------------------------------------- host-side layout
enum class Layout { None, Tkf, Mix11, Mix21, Mix31, Mix12, Mix13 };

struct Plan {
  Layout lay = Layout::None;
  int NB = 0, NG = 0, K = 0, F = 0;
  int gmap[16];
  size_t pm_stride = 0;
};

inline bool detect(const tkfmi_problem* p, Plan& pl) {
  const int S = p->S;
  const int32_t* st = p->state_types;
  const int32_t* sg = p->state_group;
  if (S == 5 && st[0] == ST_S && st[1] == ST_M && st[2] == ST_I && st[3] == ST_D && st[4] == ST_E) {
    if (!(sg[1] == sg[2] && sg[2] == sg[3])) return false;
    pl.lay = Layout::Tkf; pl.NB = 3; pl.NG = 1; pl.gmap[0] = sg[1];
    return true;
  }
  // MixDom: S = 2 + 5KF, types SS, EE, body pattern
  static const int combos[5][2] = {{1, 1}, {2, 1}, {3, 1}, {1, 2}, {1, 3}};
  static const Layout lays[5] = {Layout::Mix11, Layout::Mix21, Layout::Mix31, Layout::Mix12, Layout::Mix13};
  for (int c = 0; c < 5; ++c) {
    const int K = combos[c][0], F = combos[c][1];
    if (S != 2 + 5 * K * F) continue;
    if (st[0] != ST_S || st[1] != ST_E) continue;
    bool ok = true;
    int gm[16];
    for (int g = 0; g < 16; ++g) gm[g] = -1;
    for (int b = 0; b < 5 * K * F && ok; ++b) {
      const int uv = (b % (5 * F)) / F;
      const int ty = uv == 0 ? ST_M : ((uv == 1 || uv == 3) ? ST_I : ST_D);
      if (st[b + 2] != ty) ok = false;
      const int grp = (b / (5 * F)) * F + (b % F);
      if (gm[grp] < 0) gm[grp] = sg[b + 2];
      else if (gm[grp] != sg[b + 2]) ok = false;  // one table per (domain, fragment)
    }
    if (!ok) continue;
    pl.lay = lays[c]; pl.NB = 5 * K * F; pl.NG = K * F; pl.K = K; pl.F = F;
    for (int g = 0; g < 16; ++g) pl.gmap[g] = gm[g] < 0 ? 0 : gm[g];
    return true;
  }
  return false;
}

inline std::string layout_name(const Plan& pl) {
  if (pl.lay == Layout::Tkf) return "LTkf";
  return "LMix<" + std::to_string(pl.K) + "," + std::to_string(pl.F) + ">";
}

inline size_t stride_for(int NB) { return (size_t)(((2 * NB * NB + 2 * NB + 1) + 15) / 16 * 16); }

inline size_t align256(size_t v) { return (v + 255) / 256 * 256; }

template <class L>
inline int launch_prep(const tkfmi_problem* p, double* pmat, hipStream_t s, std::string& err) {
  hipLaunchKernelGGL((prep_kernel<L>), dim3(p->n_trans), dim3(256), 0, s, p->log_trans, p->S,
                     p->n_trans, pmat);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("prep launch: ") + hipGetErrorString(e); return TKFMI_EHIP; }
  return TKFMI_OK;
}

// Library streams (one per launch class) and the events that fork them from
// and join them to a caller's stream.  A call holds `mu` from its start record
// to its last join, so concurrent calls never pick up each other's records.
struct Streams {
  bool init = false;
  hipStream_t key = nullptr;  // the caller stream this set serves
  hipStream_t s[4];
  hipEvent_t start, done[4];
  std::mutex mu;
};
// At most this many sets per device; further caller streams share one (their
// calls are then serialised on its class streams, still correctly ordered).
constexpr int kStreamSets = 16;
// Returns the set serving calls on `caller` on the current device, creating
// it on first use: calls on different streams get different class streams
// and overlap on the device.
int get_streams(hipStream_t caller, Streams*& out, std::string& err);

inline int nx_of(const Plan& pl) {
  return pl.lay == Layout::Tkf ? 3 : (3 * pl.K * pl.F + 1);  // nM + nD + 1
}

}  // namespace tkfmi_v2

#endif
