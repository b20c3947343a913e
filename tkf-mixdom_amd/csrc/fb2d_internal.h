// Internal interfaces between the generic (state-per-lane) and the
// cell-per-lane forward/backward implementations of libtkfmi.
#ifndef TKFMI_FB2D_INTERNAL_H
#define TKFMI_FB2D_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "tkfmi.h"

// Copies `bytes` host bytes to device `dst` on `stream` through a pinned
// staging buffer owned by the library (reused once the previous upload is
// complete).
int tkfmi_stage_upload(void* dst, const void* src, size_t bytes, hipStream_t stream,
                       std::string& err);

// Cell-per-lane path (TKF91/TKF92 and MixDom K*F <= 3 layouts).
bool tkfmi_v2_workspace(const tkfmi_problem* p, uint32_t flags, size_t* bytes);
int tkfmi_v2_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                 hipStream_t stream, uint32_t flags, std::string& err);

// Chained cell-per-lane path (same layouts, a pair's bands spread over a
// chain of workgroups).
bool tkfmi_chain_workspace(const tkfmi_problem* p, uint32_t flags, size_t* bytes);
int tkfmi_chain_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                    hipStream_t stream, uint32_t flags, std::string& err);

// Wide path (16..160 emitting states).
bool tkfmi_wide_applies(const tkfmi_problem* p);
size_t tkfmi_wide_workspace(const tkfmi_problem* p, uint32_t flags);
int tkfmi_wide_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err);

// True while this thread runs tkfmi_fb2d_batched_noalloc: uploads then copy
// from the caller's host memory (no pinned staging buffer is created) and
// every launch goes on the caller's stream (no library streams or events).
bool tkfmi_noalloc_active();

// The 16..160-state kernels' per-pair blocks: P [NBP][NBP], PT, p_start,
// p_end, p(S->E) from the log transitions (pm, n_trans blocks of
// 2 NBP^2 + 2 NBP + 8 doubles) and the probability-space emission tables
// (tab, n_tab x G x TKFMI_TAB_SZ; skipped when tab is null).
int tkfmi_wide_prep(const tkfmi_problem* p, int NB, int NBP, int e_idx, const int32_t* d_bidx, double* pm,
                    int G, double* tab, hipStream_t stream, std::string& err);

// Banded path for 16..160 emitting states (a pair's rows in bands of 16,
// one workgroup per band): applies when no pair needs checkpointing.
bool tkfmi_band_applies(const tkfmi_problem* p, uint32_t flags);
size_t tkfmi_band_workspace(const tkfmi_problem* p, uint32_t flags);
int tkfmi_band_run(const tkfmi_problem* p, const tkfmi_outputs* o, void* ws, size_t ws_bytes,
                   hipStream_t stream, uint32_t flags, std::string& err);

// Lattice bytes of one pair above which the 16..160-state kernel keeps only
// checkpoints (tkfmi_checkpoint_budget; 0 = only with TKFMI_CHECKPOINT).
size_t tkfmi_checkpoint_budget_bytes();
// Antidiagonals per checkpoint segment of the antidiagonal kernel
// (tkfmi_checkpoint_segment; 0 = K ~ sqrt(2 (Lx + Ly + 1)) per pair).
int tkfmi_checkpoint_segment_k();

// Records `msg` as this thread's tkfmi_last_error() and returns `code`.
int tkfmi_fail(int code, const std::string& msg);
// Clears this thread's error message.
void tkfmi_clear_error();

// Per-kernel timing (TKFMI_TIME_KERNELS): begin_call arms or disarms the
// recorder for this thread's call; start/stop bracket one kernel launch on
// stream `s` (start returns a slot, -1 when disarmed, -2 on error).
void tkfmi_ktime_begin_call(bool enabled);
int tkfmi_ktime_start(const char* name, int n_pairs, int64_t cells, hipStream_t s, std::string& err);
int tkfmi_ktime_stop(int slot, hipStream_t s, std::string& err);

#endif
