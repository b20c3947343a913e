// tkfmi: exp(A) of a small matrix held in LDS, computed by one whole
// workgroup (every thread calls it), for the device parameter builder
// (suffstats.hip) and the gradient path's batched exp(A) and its adjoint
// (grad_build.hip).
#ifndef TKFMI_EXPM_LDS_H
#define TKFMI_EXPM_LDS_H

#include <hip/hip_runtime.h>

namespace tkfmi_expm {

// ---- matrix exponential in LDS (whole workgroup; every thread calls it).
// The rule of jax.scipy.linalg.expm, which the reference's
// transition_matrix calls (core/ctmc.py:31-60): with a = ||Q t||_1,
// s = max(0, floor(log2(a / theta_13))) squarings of the Pade approximant
// of degree 3, 5, 7, 9 or 13 picked by a against the thresholds theta_3..9
// (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193, Table 2.3; the
// numerator coefficients b_k are that paper's), evaluated at Q t / 2^s;
// R = (V - U)^-1 (V + U) by Gauss-Jordan elimination with partial pivoting,
// then squared s times.
static __constant__ double kPadeB[5][14] = {
    {120.0, 60.0, 12.0, 1.0},
    {30240.0, 15120.0, 3360.0, 420.0, 30.0, 1.0},
    {17297280.0, 8648640.0, 1995840.0, 277200.0, 25200.0, 1512.0, 56.0, 1.0},
    {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0, 2162160.0, 110880.0,
     3960.0, 90.0, 1.0},
    {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
     129060195264000.0, 10559470521600.0, 670442572800.0, 33522128640.0, 1323241920.0,
     40840800.0, 960960.0, 16380.0, 182.0, 1.0}};
constexpr double kTheta13 = 5.371920351148152;
constexpr double kPadeThr[4] = {1.495585217958292e-2, 2.539398330063230e-1,
                                9.504178996162932e-1, 2.097847961257068e+0};
constexpr int kExpmMaxSquarings = 16;  // beyond this jax.scipy.linalg.expm returns NaN
constexpr int kExpmBufs = 8;           // A x A LDS matrices used by expm_lds

// C = X Y (A x A, row-major, all in LDS); caller syncs.
__device__ inline void mm_lds(const double* X, const double* Y, double* Cm, int A) {
  for (int q = threadIdx.x; q < A * A; q += blockDim.x) {
    const int r = q / A, c = q % A;
    double acc = 0.0;
    for (int k = 0; k < A; ++k) acc = fma(X[r * A + k], Y[k * A + c], acc);
    Cm[q] = acc;
  }
}

// out = exp(t Q) (A x A).  w: kExpmBufs * A * A doubles of LDS; sh: 4 ints.
__device__ inline void expm_lds(const double* __restrict__ Q, double t, int A, double* out, double* w,
                         int* sh) {
  const int n = A * A, tid = threadIdx.x, nt = blockDim.x;
  double* X = w;            // Q t / 2^s
  double* X2 = w + n;
  double* X4 = w + 2 * n;
  double* X6 = w + 3 * n;
  double* T1 = w + 4 * n;   // inner polynomial / X8
  double* T2 = w + 5 * n;   // U
  double* M = w + 6 * n;    // [V - U | V + U], A x 2A (T2 is dead once M is formed... kept apart)
  for (int q = tid; q < n; q += nt) X[q] = Q[q] * t;
  __syncthreads();
  if (tid == 0) {
    double nrm = 0.0;
    for (int c = 0; c < A; ++c) {
      double cs = 0.0;
      for (int r = 0; r < A; ++r) cs += fabs(X[r * A + c]);
      if (!(cs <= nrm)) nrm = cs;  // a NaN column sum is kept
    }
    int deg = 0;
    while (deg < 4 && nrm >= kPadeThr[deg]) ++deg;
    double s = floor(log2(nrm / kTheta13));
    s = (s > 0.0) ? s : 0.0;  // also 0 for nrm == 0 (log2 -> -inf)
    sh[0] = deg;
    sh[1] = (s > (double)kExpmMaxSquarings) ? -1 : (int)s;
    sh[2] = (nrm == nrm) ? 0 : 1;  // NaN input
  }
  __syncthreads();
  const int deg = sh[0], ns = sh[1];
  if (ns < 0 || sh[2]) {  // too many squarings (or NaN): NaN, as jax.scipy.linalg.expm
    for (int q = tid; q < n; q += nt) out[q] = __builtin_nan("");
    __syncthreads();
    return;
  }
  const double sc = ldexp(1.0, -ns);
  for (int q = tid; q < n; q += nt) X[q] *= sc;
  __syncthreads();
  mm_lds(X, X, X2, A);
  __syncthreads();
  if (deg >= 1) { mm_lds(X2, X2, X4, A); __syncthreads(); }
  if (deg >= 2) { mm_lds(X4, X2, X6, A); __syncthreads(); }
  const double* b = kPadeB[deg];
  // U = X * inner_odd, V = even part
  if (deg <= 3) {
    if (deg == 3) { mm_lds(X6, X2, T1, A); __syncthreads(); }  // X8
    for (int q = tid; q < n; q += nt) {
      const double id = (q / A == q % A) ? 1.0 : 0.0;
      double o = b[1] * id + b[3] * X2[q], e = b[0] * id + b[2] * X2[q];
      if (deg >= 1) { o += b[5] * X4[q]; e += b[4] * X4[q]; }
      if (deg >= 2) { o += b[7] * X6[q]; e += b[6] * X6[q]; }
      if (deg >= 3) { o += b[9] * T1[q]; e += b[8] * T1[q]; }
      M[(q / A) * 2 * A + A + q % A] = e;  // V parked in the right half
      X4[q] = o;                           // odd inner polynomial (X4 / X6 no longer needed)
    }
    __syncthreads();
    mm_lds(X, X4, T2, A);  // U
    __syncthreads();
  } else {  // degree 13
    for (int q = tid; q < n; q += nt) T1[q] = b[13] * X6[q] + b[11] * X4[q] + b[9] * X2[q];
    __syncthreads();
    mm_lds(X6, T1, T2, A);
    __syncthreads();
    for (int q = tid; q < n; q += nt) {
      const double id = (q / A == q % A) ? 1.0 : 0.0;
      T2[q] += b[7] * X6[q] + b[5] * X4[q] + b[3] * X2[q] + b[1] * id;
      T1[q] = b[12] * X6[q] + b[10] * X4[q] + b[8] * X2[q];
    }
    __syncthreads();
    mm_lds(X6, T1, M + 0, A);  // A6 * inner_even into a temporary (row stride A)
    __syncthreads();
    for (int q = tid; q < n; q += nt) {
      const double id = (q / A == q % A) ? 1.0 : 0.0;
      T1[q] = M[q] + b[6] * X6[q] + b[4] * X4[q] + b[2] * X2[q] + b[0] * id;  // V
    }
    __syncthreads();
    for (int q = tid; q < n; q += nt) M[(q / A) * 2 * A + A + q % A] = T1[q];
    __syncthreads();
    mm_lds(X, T2, X4, A);  // U
    __syncthreads();
    for (int q = tid; q < n; q += nt) T2[q] = X4[q];
    __syncthreads();
  }
  // M = [V - U | V + U]
  for (int q = tid; q < n; q += nt) {
    const int r = q / A, c = q % A;
    const double v = M[r * 2 * A + A + c], u = T2[q];
    M[r * 2 * A + c] = v - u;
    M[r * 2 * A + A + c] = v + u;
  }
  __syncthreads();
  double* fac = T1;  // per-row elimination factors
  for (int k = 0; k < A; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(M[k * 2 * A + k]);
      for (int r = k + 1; r < A; ++r) {
        const double v = fabs(M[r * 2 * A + k]);
        if (v > best) { best = v; p = r; }
      }
      sh[3] = p;
    }
    __syncthreads();
    const int p = sh[3];
    if (p != k) {
      for (int c = tid; c < 2 * A; c += nt) {
        const double tmp = M[k * 2 * A + c];
        M[k * 2 * A + c] = M[p * 2 * A + c];
        M[p * 2 * A + c] = tmp;
      }
      __syncthreads();
    }
    if (tid < A) fac[tid] = (tid == k) ? 0.0 : M[tid * 2 * A + k] / M[k * 2 * A + k];
    __syncthreads();
    for (int q = tid; q < A * 2 * A; q += nt) {
      const int r = q / (2 * A), c = q % (2 * A);
      if (r != k && c > k) M[q] = fma(-fac[r], M[k * 2 * A + c], M[q]);
    }
    __syncthreads();
  }
  double* R = X2;
  for (int q = tid; q < n; q += nt) {
    const int r = q / A, c = q % A;
    R[q] = M[r * 2 * A + A + c] / M[r * 2 * A + r];
  }
  __syncthreads();
  for (int i = 0; i < ns; ++i) {
    double* Rn = (R == X2) ? X4 : X2;
    mm_lds(R, R, Rn, A);
    __syncthreads();
    R = Rn;
  }
  for (int q = tid; q < n; q += nt) out[q] = R[q];
  __syncthreads();
}

}  // namespace tkfmi_expm

#endif
