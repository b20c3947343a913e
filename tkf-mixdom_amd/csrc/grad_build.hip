// tkfmi: native forward and adjoint of the parameter builders on the
// gradient path (tkfmi/vjp.py), so that a value + gradient step runs a few
// kernels where the torch builders launch over a thousand tiny ones:
//
//  * TKF91 / conditioned TKF91 / TKF92 transition matrices and their
//    Jacobians with respect to (lambda, mu, t, ext), in forward-mode dual
//    arithmetic (one thread per parameter set).  The closed forms, guard
//    constants and the lambda ~ mu switch are those of core/bdi.py:40-97 and
//    core/params.py:29-108, in the order tkfmi/params.py evaluates them, and
//    each guarded branch differentiates as torch.where / torch.clamp do.
//  * The collapsed MixDom chi = T_KK + T_KZ (I - T_ZZ)^-1 T_ZK from the
//    exploded machine's atoms (every transition is the product of two atoms,
//    tkfmi/mixdom.py) and its adjoint: with W = (I - T_ZZ)^-1 T_ZK and
//    Y = (I - T_ZZ)^-T T_KZ^T G, the gradients are G (T_KK), G W^T (T_KZ),
//    Y (T_ZK) and Y W^T (T_ZZ), then through each transition's two atoms.
//    I - T_ZZ is a nonsingular M-matrix (the null states reach End), so the
//    Gauss-Jordan solves need no pivoting.  One workgroup per pair.
//  * exp(A) of a batch of small matrices (the rule of jax.scipy.linalg.expm,
//    core/ctmc.py:31-60) and its adjoint, the Frechet derivative L(A^T, G)
//    of exp at A^T in the direction G.  L is computed alongside the
//    degree-13 Pade approximant and its squarings, each product X Y paired
//    with its derivative X dY + dX Y (Al-Mohy and Higham, "Computing the
//    Frechet derivative of the matrix exponential", SIAM J. Matrix Anal.
//    Appl. 30 (2009) 1639-1657, alg. 6.4, with degree 13 throughout):
//    n x n products only, where exp of the doubled matrix
//    [[A^T, G], [0, A^T]] would take eight times the work.  One workgroup per
//    matrix, everything in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "expm_lds.h"
#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_gb {

using namespace tkfmi_expm;

// ------------------------------------------------------------- dual numbers
// v and its partial derivatives with respect to (lambda, mu, t, ext).
struct D4 {
  double v, d[4];
};
__device__ inline D4 cst(double v) { return {v, {0.0, 0.0, 0.0, 0.0}}; }
__device__ inline D4 var(double v, int k) {
  D4 r = cst(v);
  r.d[k] = 1.0;
  return r;
}
__device__ inline D4 operator+(const D4& a, const D4& b) {
  return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2], a.d[3] + b.d[3]}};
}
__device__ inline D4 operator-(const D4& a, const D4& b) {
  return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2], a.d[3] - b.d[3]}};
}
__device__ inline D4 operator-(const D4& a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2], -a.d[3]}}; }
__device__ inline D4 operator*(const D4& a, const D4& b) {
  D4 r;
  r.v = a.v * b.v;
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ inline D4 operator/(const D4& a, const D4& b) {
  D4 r;
  r.v = a.v / b.v;
  for (int k = 0; k < 4; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
__device__ inline D4 operator+(double a, const D4& b) { return cst(a) + b; }
__device__ inline D4 operator-(double a, const D4& b) { return cst(a) - b; }
__device__ inline D4 operator*(double a, const D4& b) { return cst(a) * b; }
__device__ inline D4 operator/(const D4& a, double b) { return a / cst(b); }
__device__ inline D4 dexp(const D4& a) {
  const double e = exp(a.v);
  return {e, {e * a.d[0], e * a.d[1], e * a.d[2], e * a.d[3]}};
}
// torch.clamp(x, min=m): x where x >= m (the gradient passes), else the constant m
__device__ inline D4 clamp_min(const D4& x, double m) { return (x.v >= m) ? x : cst(m); }

constexpr double kEqualRate = 1e-4;  // |1 - kappa| below which the lambda = mu limits apply

// T (row-major 5 x 5, states S M I D E) of TKF91 (mode 0), conditioned TKF91
// (mode 1) or TKF92 (mode 2, ext the fragment extension probability).
__device__ void tkf_trans_dual(D4 lam, D4 mu, D4 t, D4 ext, int mode, D4 (&T)[25]) {
  const D4 alpha = dexp(-(mu * t));
  // beta (core/bdi.py; tkfmi/params.py tkf_beta)
  const D4 s = mu * t;
  const D4 eta = dexp(-(lam * t));
  const D4 al = dexp(-s);
  const D4 numer = lam * (eta - al);
  const D4 denom = mu * eta - lam * al;
  const bool small = fabs(denom.v) < 1e-30;
  const D4 general_b = small ? cst(0.0) : numer / denom;
  const D4 limit_b = s / (1.0 + s);
  const D4 kappa = lam / mu;
  const bool equal = fabs(1.0 - kappa.v) < kEqualRate;
  const D4 beta = equal ? limit_b : general_b;
  // gamma (tkf_gamma)
  const D4 Phi = 1.0 - al;
  const D4 general_g = (fabs(Phi.v) < 1e-30) ? cst(0.0) : 1.0 - mu * beta / (lam * clamp_min(Phi, 1e-30));
  const D4 phi = (s.v < 1e-10) ? 1.0 - s / 2.0 : Phi / clamp_min(s, 1e-30);
  const D4 limit_g = 1.0 - cst(1.0) / ((1.0 + s) * phi);
  const D4 gamma = equal ? limit_g : general_g;
  // rows S, M, I use beta, row D gamma (core/params.py:29-89)
  const bool cond = mode == 1;
  const D4 kk = cond ? cst(1.0) : kappa;
  const D4 e_b = cond ? (1.0 - beta) : (1.0 - beta) * (1.0 - kappa);
  const D4 e_g = cond ? (1.0 - gamma) : (1.0 - gamma) * (1.0 - kappa);
  const D4 rb[5] = {cst(0.0), (1.0 - beta) * kk * alpha, beta, (1.0 - beta) * kk * (1.0 - alpha), e_b};
  const D4 rd[5] = {cst(0.0), (1.0 - gamma) * kk * alpha, gamma, (1.0 - gamma) * kk * (1.0 - alpha), e_g};
  for (int c = 0; c < 5; ++c) {
    T[c] = rb[c];
    T[5 + c] = rb[c];
    T[10 + c] = rb[c];
    T[15 + c] = rd[c];
    T[20 + c] = cst(0.0);
  }
  if (mode == 2) {  // TKF92: rows M, I, D scaled by 1 - ext, ext added on the diagonal (core/params.py:92-108)
    const D4 keep = 1.0 - ext;
    for (int r = 1; r <= 3; ++r)
      for (int c = 0; c < 5; ++c) {
        T[5 * r + c] = T[5 * r + c] * keep;
        T[5 * r + c] = T[5 * r + c] + ((r == c) ? ext : cst(0.0));
      }
  }
}

__global__ void tkf_jac_kernel(const double* __restrict__ lam, const double* __restrict__ mu,
                               const double* __restrict__ t, const double* __restrict__ ext, int64_t n,
                               int mode, double* __restrict__ T, double* __restrict__ J) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    D4 M[25];
    tkf_trans_dual(var(lam[q], 0), var(mu[q], 1), var(t[q], 2), ext ? var(ext[q], 3) : cst(0.0), mode, M);
    for (int i = 0; i < 25; ++i) {
      T[q * 25 + i] = M[i].v;
      for (int k = 0; k < 4; ++k) J[(q * 25 + i) * 4 + k] = M[i].d[k];
    }
  }
}

// ------------------------------------------------------- chi by elimination
struct ChiArgs {
  const double* atoms;   // [B, n_atoms]
  const int32_t* a1;     // [n_e] atom indices of each transition's two factors
  const int32_t* a2;
  const int32_t* flat;   // [n_e] r * N0 + c
  const int32_t* kmap;   // [N0] index among the kept states, or -1
  const int32_t* zmap;   // [N0] index among the eliminated states, or -1
  int32_t n_atoms, n_e, N0, nk, nz;
  double* ws;            // per pair: nk*nk + nk*nz + nz*(nz + nk) doubles
};

__device__ inline size_t chi_slice(const ChiArgs& a) {
  return (size_t)a.nk * a.nk + (size_t)a.nk * a.nz + (size_t)a.nz * (a.nz + a.nk);
}

// Gauss-Jordan on the nz x (nz + nk) matrix M (no pivoting; whole
// workgroup): afterwards the right block divided by the diagonal is the
// solution.  fac: nz doubles.
__device__ void gauss_jordan(double* M, int nz, int nk, double* fac) {
  const int ld = nz + nk, tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < nz; ++k) {
    const double piv = M[(size_t)k * ld + k];
    for (int r = tid; r < nz; r += nt) fac[r] = (r == k) ? 0.0 : M[(size_t)r * ld + k] / piv;
    __syncthreads();
    const int w = ld - k - 1;
    for (int q = tid; q < nz * w; q += nt) {
      const int r = q / w, c = k + 1 + q % w;
      if (r != k) M[(size_t)r * ld + c] = fma(-fac[r], M[(size_t)k * ld + c], M[(size_t)r * ld + c]);
    }
    __syncthreads();
  }
}

// The solves' nz x (nz + nk) matrix lives in LDS when it fits
// (kChiLdsDoubles), else in the pair's workspace slice.
constexpr int kChiLdsDoubles = 14 * 1024;

// chi [B, nk, nk] and W = (I - T_ZZ)^-1 T_ZK [B, nz, nk].
__global__ __launch_bounds__(256) void chi_fwd_kernel(ChiArgs a, double* __restrict__ chi, double* __restrict__ Wout,
                                                      int lds_m) {
  extern __shared__ double s_dyn[];
  double* s_fac = s_dyn;
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int nk = a.nk, nz = a.nz, ld = nz + nk;
  double* Tkk = a.ws + (size_t)b * chi_slice(a);
  double* Tkz = Tkk + (size_t)nk * nk;
  double* Mz = lds_m ? s_dyn + nz : Tkz + (size_t)nk * nz;  // [I - T_ZZ | T_ZK]
  const double* at = a.atoms + (size_t)b * a.n_atoms;
  for (size_t q = tid; q < (size_t)nk * nk + (size_t)nk * nz; q += nt) Tkk[q] = 0.0;
  for (int q = tid; q < nz * ld; q += nt) Mz[q] = 0.0;
  __syncthreads();
  for (int z = tid; z < nz; z += nt) Mz[(size_t)z * ld + z] = 1.0;
  __syncthreads();
  for (int e = tid; e < a.n_e; e += nt) {
    const double v = at[a.a1[e]] * at[a.a2[e]];
    const int r = a.flat[e] / a.N0, c = a.flat[e] % a.N0;
    const int kr = a.kmap[r], kc = a.kmap[c], zr = a.zmap[r], zc = a.zmap[c];
    if (kr >= 0 && kc >= 0) Tkk[kr * nk + kc] = v;
    else if (kr >= 0) Tkz[(size_t)kr * nz + zc] = v;
    else if (kc >= 0) Mz[(size_t)zr * ld + nz + kc] = v;
    else Mz[(size_t)zr * ld + zc] = (zr == zc) ? 1.0 - v : -v;
  }
  __syncthreads();
  gauss_jordan(Mz, nz, nk, s_fac);
  double* W = Wout + (size_t)b * nz * nk;
  for (int q = tid; q < nz * nk; q += nt) {
    const int z = q / nk, j = q % nk;
    W[q] = Mz[(size_t)z * ld + nz + j] / Mz[(size_t)z * ld + z];
  }
  __syncthreads();
  double* out = chi + (size_t)b * nk * nk;
  for (int q = tid; q < nk * nk; q += nt) {
    const int i = q / nk, j = q % nk;
    double acc = 0.0;
    for (int z = 0; z < nz; ++z) acc = fma(Tkz[(size_t)i * nz + z], W[(size_t)z * nk + j], acc);
    out[q] = Tkk[q] + acc;
  }
}

// d loss / d atoms [B, n_atoms] from G = d loss / d chi [B, nk, nk] and the
// forward's W.
__global__ __launch_bounds__(256) void chi_bwd_kernel(ChiArgs a, const double* __restrict__ G,
                                                      const double* __restrict__ Wall, double* __restrict__ gatoms,
                                                      int lds_m) {
  extern __shared__ double s_dyn[];
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int nk = a.nk, nz = a.nz, ld = nz + nk;
  double* s_fac = s_dyn;             // nz
  double* s_ga = s_dyn + nz;         // n_atoms
  double* Tkz = a.ws + (size_t)b * chi_slice(a);
  double* Mt = lds_m ? s_ga + a.n_atoms : Tkz + (size_t)nk * nz;  // [(I - T_ZZ)^T | T_KZ^T G]
  const double* at = a.atoms + (size_t)b * a.n_atoms;
  const double* Gb = G + (size_t)b * nk * nk;
  const double* W = Wall + (size_t)b * nz * nk;
  for (size_t q = tid; q < (size_t)nk * nz; q += nt) Tkz[q] = 0.0;
  for (int q = tid; q < nz * ld; q += nt) Mt[q] = 0.0;
  for (int q = tid; q < a.n_atoms; q += nt) s_ga[q] = 0.0;
  __syncthreads();
  for (int z = tid; z < nz; z += nt) Mt[(size_t)z * ld + z] = 1.0;
  __syncthreads();
  for (int e = tid; e < a.n_e; e += nt) {
    const int r = a.flat[e] / a.N0, c = a.flat[e] % a.N0;
    const int kr = a.kmap[r], zc = a.zmap[c], zr = a.zmap[r];
    if (kr >= 0 && zc >= 0) Tkz[(size_t)kr * nz + zc] = at[a.a1[e]] * at[a.a2[e]];
    else if (zr >= 0 && zc >= 0) {
      const double v = at[a.a1[e]] * at[a.a2[e]];
      Mt[(size_t)zc * ld + zr] = (zr == zc) ? 1.0 - v : -v;  // transposed
    }
  }
  __syncthreads();
  for (int q = tid; q < nz * nk; q += nt) {  // right block: T_KZ^T G
    const int z = q / nk, j = q % nk;
    double acc = 0.0;
    for (int i = 0; i < nk; ++i) acc = fma(Tkz[(size_t)i * nz + z], Gb[i * nk + j], acc);
    Mt[(size_t)z * ld + nz + j] = acc;
  }
  __syncthreads();
  gauss_jordan(Mt, nz, nk, s_fac);
  for (int q = tid; q < nz * nk; q += nt) {  // Y in place of the right block
    const int z = q / nk, j = q % nk;
    Mt[(size_t)z * ld + nz + j] /= Mt[(size_t)z * ld + z];
  }
  __syncthreads();
  for (int e = tid; e < a.n_e; e += nt) {
    const int r = a.flat[e] / a.N0, c = a.flat[e] % a.N0;
    const int kr = a.kmap[r], kc = a.kmap[c], zr = a.zmap[r], zc = a.zmap[c];
    double g;
    if (kr >= 0 && kc >= 0) g = Gb[kr * nk + kc];
    else if (kr >= 0) {  // (G W^T)[kr, zc]
      g = 0.0;
      for (int j = 0; j < nk; ++j) g = fma(Gb[kr * nk + j], W[(size_t)zc * nk + j], g);
    } else if (kc >= 0) g = Mt[(size_t)zr * ld + nz + kc];  // Y[zr, kc]
    else {  // (Y W^T)[zr, zc]
      g = 0.0;
      for (int j = 0; j < nk; ++j) g = fma(Mt[(size_t)zr * ld + nz + j], W[(size_t)zc * nk + j], g);
    }
    atomicAdd(s_ga + a.a1[e], g * at[a.a2[e]]);
    atomicAdd(s_ga + a.a2[e], g * at[a.a1[e]]);
  }
  __syncthreads();
  for (int q = tid; q < a.n_atoms; q += nt) gatoms[(size_t)b * a.n_atoms + q] = s_ga[q];
}

// --------------------------------------------------------------------- expm
constexpr int kExpmThreads = 256;

size_t expm_lds_bytes(int m) { return (size_t)(2 + kExpmBufs) * m * m * 8 + 64; }

// out[b] = exp(A[b]) (n x n).
__global__ __launch_bounds__(kExpmThreads) void expm_fwd_kernel(const double* __restrict__ A, int n,
                                                                 double* __restrict__ out) {
  extern __shared__ double s_dyn[];
  const int m2 = n * n;
  double* s_in = s_dyn;
  double* s_out = s_in + m2;
  double* s_w = s_out + m2;
  int* s_sh = reinterpret_cast<int*>(s_w + kExpmBufs * m2);
  const double* Ab = A + (size_t)blockIdx.x * m2;
  for (int q = threadIdx.x; q < m2; q += blockDim.x) s_in[q] = Ab[q];
  __syncthreads();
  expm_lds(s_in, 1.0, n, s_out, s_w, s_sh);
  double* ob = out + (size_t)blockIdx.x * m2;
  for (int q = threadIdx.x; q < m2; q += blockDim.x) ob[q] = s_out[q];
}

// C (+)= X Y (n x n, row-major, LDS); caller syncs.
__device__ inline void mm_to(const double* X, const double* Y, double* C, int n, bool add) {
  for (int q = threadIdx.x; q < n * n; q += blockDim.x) {
    const int r = q / n, c = q % n;
    double acc = add ? C[q] : 0.0;
    for (int k = 0; k < n; ++k) acc = fma(X[r * n + k], Y[k * n + c], acc);
    C[q] = acc;
  }
}

constexpr int kFrechetBufs = 18;  // n x n LDS matrices used by expm_vjp_kernel
size_t frechet_lds_bytes(int n) { return (size_t)kFrechetBufs * n * n * 8 + (size_t)n * 8 + 64; }

// out[b] = L(A[b]^T, G[b]), the adjoint of exp's Frechet derivative at A[b]:
// with X = A^T / 2^s and E = G / 2^s (s squarings as exp(A) takes them),
// the degree-13 Pade approximant R = (V - U)^-1 (V + U) and its derivative
// L = (V - U)^-1 ((LU + LV) + (LU - LV) R), then s times L <- R L + L R,
// R <- R R.
__global__ __launch_bounds__(kExpmThreads) void expm_vjp_kernel(const double* __restrict__ A,
                                                                 const double* __restrict__ G, int n,
                                                                 double* __restrict__ out) {
  extern __shared__ double s_dyn[];
  const int n2 = n * n, tid = threadIdx.x, nt = blockDim.x;
  double* X = s_dyn;
  double* E = X + n2;
  double* X2 = E + n2;
  double* L2 = X2 + n2;
  double* X4 = L2 + n2;
  double* L4 = X4 + n2;
  double* X6 = L4 + n2;
  double* L6 = X6 + n2;
  double* T = L6 + n2;
  double* LT = T + n2;
  double* W = LT + n2;
  double* LW = W + n2;
  double* U = LW + n2;
  double* LU = U + n2;
  double* M = LU + n2;  // n x 4n: [V - U | V + U | LU + LV | LU - LV]
  double* fac = M + 4 * n2;
  __shared__ int s_ns, s_piv;
  const double* Ab = A + (size_t)blockIdx.x * n2;
  const double* Gb = G + (size_t)blockIdx.x * n2;
  double* ob = out + (size_t)blockIdx.x * n2;
  for (int q = tid; q < n2; q += nt) {
    const int r = q / n, c = q % n;
    X[q] = Ab[c * n + r];
    E[q] = Gb[q];
  }
  __syncthreads();
  if (tid == 0) {  // the squarings exp(A) takes (expm_lds: from the 1-norm of A)
    double nrm = 0.0;
    for (int c = 0; c < n; ++c) {
      double cs = 0.0;
      for (int r = 0; r < n; ++r) cs += fabs(Ab[r * n + c]);
      if (!(cs <= nrm)) nrm = cs;
    }
    double sq = floor(log2(nrm / kTheta13));
    sq = (sq > 0.0) ? sq : 0.0;
    s_ns = (nrm != nrm || sq > (double)kExpmMaxSquarings) ? -1 : (int)sq;
  }
  __syncthreads();
  const int ns = s_ns;
  if (ns < 0) {  // exp(A) is NaN there (expm_lds): so is its gradient
    for (int q = tid; q < n2; q += nt) ob[q] = __builtin_nan("");
    return;
  }
  const double sc = ldexp(1.0, -ns);
  for (int q = tid; q < n2; q += nt) { X[q] *= sc; E[q] *= sc; }
  __syncthreads();
  const double* b = kPadeB[4];
  mm_to(X, X, X2, n, false);
  mm_to(X, E, L2, n, false);
  __syncthreads();
  mm_to(E, X, L2, n, true);
  mm_to(X2, X2, X4, n, false);
  __syncthreads();
  mm_to(X2, L2, L4, n, false);
  mm_to(X4, X2, X6, n, false);
  __syncthreads();
  mm_to(L2, X2, L4, n, true);
  mm_to(X4, L2, L6, n, false);
  __syncthreads();
  mm_to(L4, X2, L6, n, true);
  __syncthreads();
  // odd part: W = X6 (b13 X6 + b11 X4 + b9 X2) + b7 X6 + b5 X4 + b3 X2 + b1 I, U = X W
  for (int q = tid; q < n2; q += nt) {
    T[q] = b[13] * X6[q] + b[11] * X4[q] + b[9] * X2[q];
    LT[q] = b[13] * L6[q] + b[11] * L4[q] + b[9] * L2[q];
  }
  __syncthreads();
  mm_to(X6, T, W, n, false);
  mm_to(L6, T, LW, n, false);
  __syncthreads();
  mm_to(X6, LT, LW, n, true);
  for (int q = tid; q < n2; q += nt) {
    const double id = (q / n == q % n) ? 1.0 : 0.0;
    W[q] += b[7] * X6[q] + b[5] * X4[q] + b[3] * X2[q] + b[1] * id;
  }
  __syncthreads();
  for (int q = tid; q < n2; q += nt) LW[q] += b[7] * L6[q] + b[5] * L4[q] + b[3] * L2[q];
  mm_to(X, W, U, n, false);
  mm_to(E, W, LU, n, false);
  __syncthreads();
  mm_to(X, LW, LU, n, true);
  // even part: V = X6 (b12 X6 + b10 X4 + b8 X2) + b6 X6 + b4 X4 + b2 X2 + b0 I (into W, LW)
  for (int q = tid; q < n2; q += nt) {
    T[q] = b[12] * X6[q] + b[10] * X4[q] + b[8] * X2[q];
    LT[q] = b[12] * L6[q] + b[10] * L4[q] + b[8] * L2[q];
  }
  __syncthreads();
  mm_to(X6, T, W, n, false);
  mm_to(L6, T, LW, n, false);
  __syncthreads();
  mm_to(X6, LT, LW, n, true);
  for (int q = tid; q < n2; q += nt) {
    const double id = (q / n == q % n) ? 1.0 : 0.0;
    W[q] += b[6] * X6[q] + b[4] * X4[q] + b[2] * X2[q] + b[0] * id;
  }
  __syncthreads();
  const int ld = 4 * n;
  for (int q = tid; q < n2; q += nt) {
    LW[q] += b[6] * L6[q] + b[4] * L4[q] + b[2] * L2[q];
  }
  __syncthreads();
  for (int q = tid; q < n2; q += nt) {
    const int r = q / n, c = q % n;
    M[r * ld + c] = W[q] - U[q];
    M[r * ld + n + c] = W[q] + U[q];
    M[r * ld + 2 * n + c] = LU[q] + LW[q];
    M[r * ld + 3 * n + c] = LU[q] - LW[q];
  }
  __syncthreads();
  // Gauss-Jordan with partial pivoting on M (the rule expm_lds solves by)
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(M[k * ld + k]);
      for (int r = k + 1; r < n; ++r) {
        const double v = fabs(M[r * ld + k]);
        if (v > best) { best = v; p = r; }
      }
      s_piv = p;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k) {
      for (int c = tid; c < ld; c += nt) {
        const double tmp = M[k * ld + c];
        M[k * ld + c] = M[p * ld + c];
        M[p * ld + c] = tmp;
      }
      __syncthreads();
    }
    for (int r = tid; r < n; r += nt) fac[r] = (r == k) ? 0.0 : M[r * ld + k] / M[k * ld + k];
    __syncthreads();
    const int w = ld - k - 1;
    for (int q = tid; q < n * w; q += nt) {
      const int r = q / w, c = k + 1 + q % w;
      if (r != k) M[r * ld + c] = fma(-fac[r], M[k * ld + c], M[r * ld + c]);
    }
    __syncthreads();
  }
  double* R = X2;   // (V - U)^-1 (V + U)
  double* P = X4;   // (V - U)^-1 (LU + LV)
  double* Q = X6;   // (V - U)^-1 (LU - LV)
  for (int q = tid; q < n2; q += nt) {
    const int r = q / n, c = q % n;
    const double d = M[r * ld + r];
    R[q] = M[r * ld + n + c] / d;
    P[q] = M[r * ld + 2 * n + c] / d;
    Q[q] = M[r * ld + 3 * n + c] / d;
  }
  __syncthreads();
  double* L = L2;
  for (int q = tid; q < n2; q += nt) L[q] = P[q];
  __syncthreads();
  mm_to(Q, R, L, n, true);  // L = P + Q R
  __syncthreads();
  double* R2 = T;
  double* Ln = LT;
  for (int i = 0; i < ns; ++i) {  // L <- R L + L R, R <- R R
    mm_to(R, L, Ln, n, false);
    mm_to(R, R, R2, n, false);
    __syncthreads();
    mm_to(L, R, Ln, n, true);
    __syncthreads();
    double* t1 = R; R = R2; R2 = t1;
    double* t2 = L; L = Ln; Ln = t2;
  }
  for (int q = tid; q < n2; q += nt) ob[q] = L[q];
}

}  // namespace tkfmi_gb

using namespace tkfmi_gb;

namespace {
int launch_check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return TKFMI_OK;
}

int chi_args(ChiArgs& a, const double* atoms, int32_t n_atoms, const int32_t* a1, const int32_t* a2,
             const int32_t* flat, int32_t n_e, const int32_t* kmap, const int32_t* zmap, int32_t N0, int32_t nk,
             int32_t nz, int64_t B, void* ws, size_t ws_bytes) {
  if (B < 0 || n_atoms <= 0 || n_e < 0 || N0 <= 0 || nk <= 0 || nz < 0 || nk + nz > N0)
    return tkfmi_fail(TKFMI_EINVAL, "chi: bad sizes");
  if (!atoms || !a1 || !a2 || !flat || !kmap || !zmap) return tkfmi_fail(TKFMI_EINVAL, "chi: pointer missing");
  const size_t need = (size_t)B * ((size_t)nk * nk + (size_t)nk * nz + (size_t)nz * (nz + nk)) * 8;
  if (!ws || ws_bytes < need) return tkfmi_fail(TKFMI_EWORKSPACE, "chi: workspace too small: need " + std::to_string(need));
  a = ChiArgs{atoms, a1, a2, flat, kmap, zmap, n_atoms, n_e, N0, nk, nz, static_cast<double*>(ws)};
  return TKFMI_OK;
}
}  // namespace

extern "C" {

int tkfmi_tkf_trans_jac(const double* lam, const double* mu, const double* t, const double* ext, int64_t n,
                        int32_t mode, double* T, double* J, void* stream_v) {
  tkfmi_clear_error();
  if (n < 0 || mode < 0 || mode > 2) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_tkf_trans_jac: n < 0 or mode not 0..2");
  if (n == 0) return TKFMI_OK;
  if (!lam || !mu || !t || !T || !J || (mode == 2 && !ext))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_tkf_trans_jac: pointer missing");
  const int blocks = (int)std::min<int64_t>((n + 127) / 128, 4096);
  hipLaunchKernelGGL(tkf_jac_kernel, dim3(blocks), dim3(128), 0, reinterpret_cast<hipStream_t>(stream_v), lam, mu,
                     t, mode == 2 ? ext : nullptr, n, mode, T, J);
  return launch_check("tkf_jac_kernel");
}

size_t tkfmi_chi_workspace_bytes(int64_t B, int32_t nk, int32_t nz) {
  if (B < 0 || nk < 0 || nz < 0) return 0;
  return (size_t)B * ((size_t)nk * nk + (size_t)nk * nz + (size_t)nz * (nz + nk)) * 8;
}

int tkfmi_chi_forward(const double* atoms, int64_t B, int32_t n_atoms, const int32_t* a1, const int32_t* a2,
                      const int32_t* flat, int32_t n_e, const int32_t* kmap, const int32_t* zmap, int32_t N0,
                      int32_t nk, int32_t nz, double* chi, double* W, void* ws, size_t ws_bytes, void* stream_v) {
  tkfmi_clear_error();
  ChiArgs a;
  int rc = chi_args(a, atoms, n_atoms, a1, a2, flat, n_e, kmap, zmap, N0, nk, nz, B, ws, ws_bytes);
  if (rc != TKFMI_OK) return rc;
  if (!chi || !W) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_chi_forward: output pointer missing");
  if (B == 0) return TKFMI_OK;
  const int lds_m = (size_t)nz * (nz + nk) <= (size_t)kChiLdsDoubles;
  const size_t lds = (size_t)(std::max(nz, 1) + (lds_m ? nz * (nz + nk) : 0)) * 8;
  hipError_t e = hipFuncSetAttribute((const void*)chi_fwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(chi_fwd_kernel, dim3((unsigned)B), dim3(256), lds, reinterpret_cast<hipStream_t>(stream_v), a,
                     chi, W, lds_m);
  return launch_check("chi_fwd_kernel");
}

int tkfmi_chi_backward(const double* atoms, int64_t B, int32_t n_atoms, const int32_t* a1, const int32_t* a2,
                       const int32_t* flat, int32_t n_e, const int32_t* kmap, const int32_t* zmap, int32_t N0,
                       int32_t nk, int32_t nz, const double* G, const double* W, double* gatoms, void* ws,
                       size_t ws_bytes, void* stream_v) {
  tkfmi_clear_error();
  ChiArgs a;
  int rc = chi_args(a, atoms, n_atoms, a1, a2, flat, n_e, kmap, zmap, N0, nk, nz, B, ws, ws_bytes);
  if (rc != TKFMI_OK) return rc;
  if (!G || !W || !gatoms) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_chi_backward: pointer missing");
  if (nz + n_atoms > 8192) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_chi_backward: nz + n_atoms > 8192");
  if (B == 0) return TKFMI_OK;
  const int lds_m = (size_t)nz * (nz + nk) + nz + n_atoms <= (size_t)kChiLdsDoubles;
  const size_t lds = (size_t)(nz + n_atoms + (lds_m ? nz * (nz + nk) : 0)) * 8;
  hipError_t e = hipFuncSetAttribute((const void*)chi_bwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(chi_bwd_kernel, dim3((unsigned)B), dim3(256), lds, reinterpret_cast<hipStream_t>(stream_v), a, G,
                     W, gatoms, lds_m);
  return launch_check("chi_bwd_kernel");
}

int tkfmi_expm_batched(const double* A, int64_t count, int32_t n, double* out, void* stream_v) {
  tkfmi_clear_error();
  if (count < 0 || n < 1 || n > TKFMI_EXPM_MAX_N) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_expm_batched: n must be 1..40");
  if (count == 0) return TKFMI_OK;
  if (!A || !out) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_expm_batched: pointer missing");
  const size_t lds = expm_lds_bytes(n);
  hipError_t e = hipFuncSetAttribute((const void*)expm_fwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(expm_fwd_kernel, dim3((unsigned)count), dim3(kExpmThreads), lds,
                     reinterpret_cast<hipStream_t>(stream_v), A, n, out);
  return launch_check("expm_fwd_kernel");
}

int tkfmi_expm_vjp_batched(const double* A, const double* G, int64_t count, int32_t n, double* out, void* stream_v) {
  tkfmi_clear_error();
  if (count < 0 || n < 1 || 2 * n > TKFMI_EXPM_MAX_N)
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_expm_vjp_batched: n must be 1..20");
  if (count == 0) return TKFMI_OK;
  if (!A || !G || !out) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_expm_vjp_batched: pointer missing");
  const size_t lds = frechet_lds_bytes(n);
  hipError_t e = hipFuncSetAttribute((const void*)expm_vjp_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(expm_vjp_kernel, dim3((unsigned)count), dim3(kExpmThreads), lds,
                     reinterpret_cast<hipStream_t>(stream_v), A, G, n, out);
  return launch_check("expm_vjp_kernel");
}

}  // extern "C"
