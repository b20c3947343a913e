// tkfmi: device decoding of X/A/Y precompiled pairs.
//
// Reference format: python/tkfmixdom/jax/util/pair_format.py.  A pair is an
// ancestor string X and a descendant string Y over the 20 amino acids
// "ACDEFGHIKLMNPQRSTVWY" (:23-26), and its alignment A as a run-length string
// of M/I/D runs, "M5I1M3D2" (:38-90); decode_pair (:140-190) turns it into
// residue codes, the M/I/D state of every alignment column and the ancestor
// / descendant character of each column, which the training loop streams
// from its precompiled shards (python/train_pfam.py:984-1075).
//
// Here the caller uploads the raw text once (X, Y and A strings
// concatenated) and one launch decodes it in HBM, one workgroup per pair:
//   * X and Y bytes -> codes 0..19 (an invalid character is flagged and
//     written as the wildcard 20);
//   * RLE tokens: an M, I or D immediately followed by one or more digits
//     (the reference's regex ([MID])(\d+); other text is skipped as its
//     finditer skips it), found with a block scan in chunks of 256
//     characters; run lengths parsed in place;
//   * exclusive scans of the run lengths give each run's first column and
//     the numbers of ancestor / descendant residues before it; the totals
//     are checked (M+D = |X|, M+I = |Y|, validate_record :193-223);
//   * the columns are written one per thread, each finding its run by
//     binary search: type 1=M 2=I 3=D, the ancestor code at M/D columns and
//     the descendant code at M/I columns (0 elsewhere).
// The residue codes feed tkfmi_fb2d_batched directly (x_cat / y_cat), the
// columns tkfmi_fb1d_batched (col_type / anc / desc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace tkfmi_pairs {

constexpr int kThr = 256;
constexpr int kMaxRun = 1 << 24;

struct PDesc {
  int64_t x_off, y_off, a_off, col_off;
  int32_t Lx, Ly, La, pad;
};

struct PParams {
  const uint8_t* x_text;
  const uint8_t* y_text;
  const uint8_t* a_text;
  const PDesc* desc;
  uint8_t* x_codes;
  uint8_t* y_codes;
  uint8_t* col_type;
  uint8_t* anc;
  uint8_t* dsc;
  int32_t* n_cols;
  int32_t* status;
  int32_t* tok_len;  // workspace [sum La]: token t of pair b at a_off[b] + t
  int32_t* tok_col;
  int32_t* tok_x;
  int32_t* tok_y;
  uint8_t* tok_type;
};

// code of an amino-acid letter A..Z in "ACDEFGHIKLMNPQRSTVWY", -1 if none
__constant__ int8_t kAA[26] = {0, -1, 1, 2, 3, 4, 5, 6, 7, -1, 8, 9, 10, 11, -1, 12, 13, 14, 15, 16, -1, 17, 18, -1, 19, -1};

__device__ inline int aa_code(unsigned c) { return (c >= 'A' && c <= 'Z') ? kAA[c - 'A'] : -1; }
__device__ inline bool is_digit(unsigned c) { return c >= '0' && c <= '9'; }
__device__ inline int run_type(unsigned c) { return c == 'M' ? 1 : (c == 'I' ? 2 : (c == 'D' ? 3 : 0)); }

// Exclusive scan of v over the 256 threads (4 waves); *total gets the sum.
__device__ inline int block_excl_scan(int v, int* s_w, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  int off = 0;
  for (int k = 0; k < w; ++k) off += s_w[k];
  *total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();  // s_w is reused by the next scan
  return off + x - v;
}

__global__ __launch_bounds__(kThr) void pairs_decode_kernel(PParams P) {
  __shared__ int s_w[4];
  __shared__ int s_flag;
  const int tid = threadIdx.x;
  const PDesc d = P.desc[blockIdx.x];
  if (tid == 0) s_flag = 0;
  __syncthreads();
  int flag = 0;
  // ---- residue codes
  for (int q = tid; q < d.Lx; q += kThr) {
    int c = aa_code(P.x_text[d.x_off + q]);
    if (c < 0) { flag |= TKFMI_PAIR_BAD_X; c = 20; }
    P.x_codes[d.x_off + q] = (uint8_t)c;
  }
  for (int q = tid; q < d.Ly; q += kThr) {
    int c = aa_code(P.y_text[d.y_off + q]);
    if (c < 0) { flag |= TKFMI_PAIR_BAD_Y; c = 20; }
    P.y_codes[d.y_off + q] = (uint8_t)c;
  }
  // ---- RLE tokens
  const uint8_t* a = P.a_text + d.a_off;
  int ntok = 0;
  for (int p0 = 0; p0 < d.La; p0 += kThr) {
    const int p = p0 + tid;
    int ty = 0;
    if (p < d.La) {
      ty = run_type(a[p]);
      if (ty && !(p + 1 < d.La && is_digit(a[p + 1]))) ty = 0;
    }
    int tot;
    const int t = ntok + block_excl_scan(ty != 0, s_w, &tot);
    if (ty) {
      int len = 0;
      for (int q = p + 1; q < d.La && is_digit(a[q]); ++q) {
        len = len * 10 + (int)(a[q] - '0');
        if (len > kMaxRun) { len = kMaxRun; flag |= TKFMI_PAIR_BAD_RUN; }
      }
      P.tok_type[d.a_off + t] = (uint8_t)ty;
      P.tok_len[d.a_off + t] = len;
    }
    ntok += tot;
  }
  __syncthreads();  // token stores visible to the whole workgroup
  // ---- run starts: columns, ancestor and descendant residues before each run
  int ncol = 0, nx = 0, ny = 0;
  for (int t0 = 0; t0 < ntok; t0 += kThr) {
    const int t = t0 + tid;
    int len = 0, ty = 0;
    if (t < ntok) { len = P.tok_len[d.a_off + t]; ty = P.tok_type[d.a_off + t]; }
    int tc, tx, tyy;
    const int ec = block_excl_scan(len, s_w, &tc);
    const int ex = block_excl_scan(ty != 2 ? len : 0, s_w, &tx);
    const int ey = block_excl_scan(ty != 3 ? len : 0, s_w, &tyy);
    if (t < ntok) {
      P.tok_col[d.a_off + t] = ncol + ec;
      P.tok_x[d.a_off + t] = nx + ex;
      P.tok_y[d.a_off + t] = ny + ey;
    }
    ncol += tc; nx += tx; ny += tyy;
  }
  if (nx != d.Lx) flag |= TKFMI_PAIR_BAD_LEN_X;
  if (ny != d.Ly) flag |= TKFMI_PAIR_BAD_LEN_Y;
  if (flag) atomicOr(&s_flag, flag);
  __syncthreads();
  const int status = s_flag;
  const bool cols_ok = !(status & (TKFMI_PAIR_BAD_LEN_X | TKFMI_PAIR_BAD_LEN_Y | TKFMI_PAIR_BAD_RUN));
  if (tid == 0) {
    P.status[blockIdx.x] = status;
    P.n_cols[blockIdx.x] = cols_ok ? ncol : 0;
  }
  if (!P.col_type || !cols_ok) return;
  // ---- columns (ncol = Lx + Ly - #M <= Lx + Ly, the pair's capacity)
  const int32_t* tcol = P.tok_col + d.a_off;
  for (int col = tid; col < ncol; col += kThr) {
    int lo = 0, hi = ntok - 1;  // last run whose first column is <= col
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tcol[mid] <= col) lo = mid; else hi = mid - 1;
    }
    const int c = col - tcol[lo];
    const int ty = P.tok_type[d.a_off + lo];
    int xa = 0, yd = 0;
    if (ty != 2) { xa = aa_code(P.x_text[d.x_off + P.tok_x[d.a_off + lo] + c]); xa = xa < 0 ? 20 : xa; }
    if (ty != 3) { yd = aa_code(P.y_text[d.y_off + P.tok_y[d.a_off + lo] + c]); yd = yd < 0 ? 20 : yd; }
    P.col_type[d.col_off + col] = (uint8_t)ty;
    P.anc[d.col_off + col] = (uint8_t)xa;
    P.dsc[d.col_off + col] = (uint8_t)yd;
  }
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

bool sizes(const tkfmi_pair_text* in, int64_t& sum_a, std::string& why) {
  if (!in || in->B < 0) { why = "null input or B < 0"; return false; }
  if (in->B > 0 && (!in->h_x_off || !in->h_x_len || !in->h_y_off || !in->h_y_len || !in->h_a_off || !in->h_a_len)) {
    why = "host offset / length arrays missing";
    return false;
  }
  sum_a = 0;
  for (int b = 0; b < in->B; ++b) {
    if (in->h_x_len[b] < 0 || in->h_y_len[b] < 0 || in->h_a_len[b] < 0) { why = "negative length"; return false; }
    if ((int64_t)in->h_x_len[b] + in->h_y_len[b] > (1 << 26)) { why = "pair too long"; return false; }
    if (in->h_x_off[b] < 0 || in->h_y_off[b] < 0 || in->h_a_off[b] < 0) { why = "negative offset"; return false; }
    sum_a = std::max<int64_t>(sum_a, in->h_a_off[b] + in->h_a_len[b]);
  }
  return true;
}

}  // namespace tkfmi_pairs

using namespace tkfmi_pairs;

extern "C" size_t tkfmi_pairs_decode_workspace_bytes(const tkfmi_pair_text* in) {
  int64_t sum_a = 0;
  std::string why;
  if (!sizes(in, sum_a, why)) return 0;
  const size_t n = (size_t)std::max<int64_t>(sum_a, 1);
  return align256(sizeof(PDesc) * (size_t)std::max(in->B, 1)) + 4 * align256(n * 4) + align256(n);
}

extern "C" int tkfmi_pairs_decode(const tkfmi_pair_text* in, const tkfmi_pair_columns* out, void* workspace,
                                  size_t workspace_bytes, void* stream_v) {
  tkfmi_clear_error();
  int64_t sum_a = 0;
  std::string why;
  if (!sizes(in, sum_a, why)) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_pairs_decode: " + why);
  if (!out || !out->x_codes || !out->y_codes || !out->n_cols || !out->status)
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_pairs_decode: output pointer missing");
  if ((out->col_type || out->anc || out->desc) && !(out->col_type && out->anc && out->desc))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_pairs_decode: col_type, anc and desc go together");
  const int B = in->B;
  if (B == 0) return TKFMI_OK;
  if (!in->x_text || !in->y_text || !in->a_text) return tkfmi_fail(TKFMI_EINVAL, "tkfmi_pairs_decode: text pointer missing");
  const size_t need = tkfmi_pairs_decode_workspace_bytes(in);
  if (!workspace || workspace_bytes < need)
    return tkfmi_fail(TKFMI_EWORKSPACE, "tkfmi_pairs_decode: workspace too small: need " + std::to_string(need));
  std::vector<PDesc> hd(B);
  int64_t col = 0;
  for (int b = 0; b < B; ++b) {
    PDesc d;
    memset(&d, 0, sizeof(d));
    d.x_off = in->h_x_off[b]; d.y_off = in->h_y_off[b]; d.a_off = in->h_a_off[b];
    d.Lx = in->h_x_len[b]; d.Ly = in->h_y_len[b]; d.La = in->h_a_len[b];
    d.col_off = col;
    col += (int64_t)d.Lx + d.Ly;
    hd[b] = d;
  }
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_v);
  char* ws = static_cast<char*>(workspace);
  std::string err;
  int rc = tkfmi_stage_upload(ws, hd.data(), sizeof(PDesc) * (size_t)B, stream, err);
  if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
  const size_t n = (size_t)std::max<int64_t>(sum_a, 1);
  size_t o = align256(sizeof(PDesc) * (size_t)B);
  PParams P;
  memset(&P, 0, sizeof(P));
  P.x_text = in->x_text; P.y_text = in->y_text; P.a_text = in->a_text;
  P.desc = reinterpret_cast<const PDesc*>(ws);
  P.tok_len = reinterpret_cast<int32_t*>(ws + o); o += align256(n * 4);
  P.tok_col = reinterpret_cast<int32_t*>(ws + o); o += align256(n * 4);
  P.tok_x = reinterpret_cast<int32_t*>(ws + o); o += align256(n * 4);
  P.tok_y = reinterpret_cast<int32_t*>(ws + o); o += align256(n * 4);
  P.tok_type = reinterpret_cast<uint8_t*>(ws + o);
  P.x_codes = out->x_codes; P.y_codes = out->y_codes;
  P.col_type = out->col_type; P.anc = out->anc; P.dsc = out->desc;
  P.n_cols = out->n_cols; P.status = out->status;
  hipLaunchKernelGGL(pairs_decode_kernel, dim3(B), dim3(kThr), 0, stream, P);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tkfmi_fail(TKFMI_EHIP, std::string("pairs decode launch: ") + hipGetErrorString(e));
  return TKFMI_OK;
}
