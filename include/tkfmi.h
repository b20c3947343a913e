/*
 * tkfmi -- MI355X (gfx950) pair-HMM forward/backward/expected-count library.
 *
 * C ABI (no C++ types, no torch types): plain pointers and sizes.  Every
 * array argument documented as "device" must be device memory on the
 * current HIP device; "host" arrays are read by the library on the CPU.
 *
 * The reference has no FFI: its operator surface is a set of Python
 * functions over FP64 JAX arrays.  The entry points below replace, for a
 * ragged batch of pairs in one launch, the reference calls named beside
 * each one (paths relative to the reference's python/ directory).  The
 * Python package tkfmi binds them with ctypes (tkfmi/_abi.py).
 *
 * Conventions
 *   - FP64 throughout; log-space inputs exactly as the reference passes
 *     them (log transition matrix, log emission tables).
 *   - Return 0 on success, a negative TKFMI_E* code on failure; the message
 *     is available from tkfmi_last_error() (thread-local).
 *   - Calls are stream-ordered and asynchronous on `stream`; numerical
 *     failure (impossible pair) is returned as a NEG_INF log-probability,
 *     not raised, as the reference does.
 */
#ifndef TKFMI_H
#define TKFMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TKFMI_ABI_VERSION 6

/* error codes */
#define TKFMI_OK 0
#define TKFMI_EINVAL (-1)       /* invalid argument / unsupported model   */
#define TKFMI_EWORKSPACE (-2)   /* workspace missing or too small         */
#define TKFMI_EHIP (-3)         /* HIP runtime error                      */

/* flags */
#define TKFMI_FORWARD_ONLY 1u    /* logP only (dp/hmm.py:1915-1916)       */
#define TKFMI_WANT_POSTERIORS 2u /* write the posterior lattice           */
#define TKFMI_NO_COUNTS 4u       /* skip expected transition counts       */
#define TKFMI_GENERIC_KERNEL 8u  /* force the generic (any-layout) kernel  */
#define TKFMI_WANT_EMIT_COUNTS 16u /* expected emission counts (see below)  */
#define TKFMI_TIME_KERNELS 32u   /* record per-kernel HIP-event timings    */
#define TKFMI_CLASS_KERNEL 64u   /* force the one-workgroup-per-pair
                                    kernels (see below)                      */
#define TKFMI_REUSE_DESCRIPTORS 128u /* the workspace already holds this
                                    problem's descriptors from an earlier
                                    call with the same problem, flags and
                                    workspace: skip their host-to-device
                                    upload, so the call is only device
                                    work (e.g. captured into a HIP graph)  */
#define TKFMI_CHECKPOINT 256u    /* checkpointing for every pair (the
                                    reference's gradient path recomputes
                                    the lattice under jax.checkpoint,
                                    python/tkfmixdom/jax/train/vjp.py
                                    remat).  Banded kernel, when the batch
                                    has enough bands: the forward keeps
                                    only its boundary rows, and each
                                    backward band recomputes its own
                                    lattice from the row above into one of
                                    as many slots as workgroups fit the
                                    device (one extra forward sweep,
                                    lattice memory one band per resident
                                    workgroup).  Otherwise the antidiagonal
                                    kernel: the forward keeps the restart
                                    state of every K-th antidiagonal, K ~
                                    sqrt(2 (Lx+Ly+1)), and the backward
                                    recomputes one K-antidiagonal segment
                                    at a time (lattice memory about
                                    3 K / (Ly+1) of the full one).  Results
                                    are identical to the full-lattice
                                    run's.  Without
                                    the flag, pairs whose lattice exceeds
                                    the budget (tkfmi_checkpoint_budget,
                                    256 MiB by default) are checkpointed.
                                    Applies to the 16..160-state kernel;
                                    the cell-per-lane and generic kernels
                                    keep their 128-B-per-cell lattice (at
                                    most 33 GB for the longest pair they
                                    accept, Lx + Ly <= 32000, which fits
                                    HBM) and ignore it                     */

/* emission tables: per group g, 21x21 match block then 21 insert then 21
 * delete entries (index 20 = wildcard), all natural-log values. */
#define TKFMI_A_PAD 21
#define TKFMI_TAB_SZ (TKFMI_A_PAD * TKFMI_A_PAD + 2 * TKFMI_A_PAD)

typedef struct tkfmi_problem {
  int32_t B;                   /* number of pairs                           */
  int32_t S;                   /* number of states                          */
  const int32_t* state_types;  /* host [S]: 0=S 1=M 2=I 3=D 4=E (core/params.py:22) */
  const int32_t* state_group;  /* host [S]: emission group of each state    */
  int32_t n_groups;            /* G                                         */
  int32_t n_trans;             /* 1 (shared) or B (per pair)                */
  const double* log_trans;     /* device [n_trans,S,S] log transitions      */
  int32_t n_tab;               /* 1 (shared) or B (per pair)                */
  const double* log_tab;       /* device [n_tab,G,TKFMI_TAB_SZ]             */
  const uint8_t* x_cat;        /* device: concatenated ancestor codes 0..20 */
  const uint8_t* y_cat;        /* device: concatenated descendant codes     */
  const int64_t* h_x_off;      /* host [B] offsets into x_cat               */
  const int64_t* h_y_off;      /* host [B] offsets into y_cat               */
  const int32_t* h_x_len;      /* host [B] Lx                               */
  const int32_t* h_y_len;      /* host [B] Ly                               */
  const double* log_emit;      /* device, or NULL (table emissions): dense
                                  log emissions, the reference's
                                  log_emit_table (dp/hmm.py:1834-1873), ragged
                                  like the posteriors -- pair b's
                                  [Lx+1][Ly+1][S] block at offset
                                  sum_{b'<b} (Lx+1)(Ly+1) S -- already cut to
                                  the real (Lx, Ly) region; -1e30 = impossible.
                                  Replaces log_tab / x_cat / y_cat (which may
                                  then be NULL); TKFMI_WANT_EMIT_COUNTS is
                                  refused (d logP / d log_emit is the posterior
                                  lattice).  Runs on the generic kernel (<= 15
                                  emitting states) or the 16..160-state kernel */
} tkfmi_problem;

typedef struct tkfmi_outputs {
  double* log_prob;   /* device [B]                                          */
  double* n_trans;    /* device [B,S,S], zero-filled by the caller, or NULL  */
  double* posteriors; /* device ragged: pair b at offset sum_{b'<b}
                         (Lx+1)(Ly+1)S, layout [i][j][state], zero-filled by
                         the caller; NULL unless TKFMI_WANT_POSTERIORS       */
  double* emit_counts; /* device [B, n_groups, TKFMI_TAB_SZ], zero-filled by
                          the caller; NULL unless TKFMI_WANT_EMIT_COUNTS.
                          Same layout as log_tab: entry (g, e) is the
                          posterior-expected number of times table entry e of
                          group g is used, i.e. d logP / d log_tab[g, e].    */
  int32_t* status;     /* device [1] or NULL, zero-filled by the caller: set
                          to 1 (never cleared by the library) when a chained
                          launch's bounded internal wait gave up, the case in
                          which the affected pairs' log_prob is NaN; lets a
                          caller tell such pairs from impossible ones (which
                          the reference filters as NaN / -inf)              */
} tkfmi_outputs;

/* Lattice bytes of one pair above which the 16..160-state kernel
 * checkpoints that pair without TKFMI_CHECKPOINT (default 256 MiB, the
 * budget of the reference's gradient configuration; 0 = never without the
 * flag).  Exception: when the banded kernel's recomputation slots would not
 * halve the selected pairs' lattices (a few long pairs) and those lattices
 * total at most 16 budgets, they are kept whole (the faster choice).
 * Process-wide; query workspace sizes after changing it.  No counterpart in
 * the reference (its remat is always on in the VJP). */
int tkfmi_checkpoint_budget(size_t bytes);

/* Antidiagonals per recomputed segment (the K of CHECKPOINT_K) when the
 * one-workgroup-per-pair antidiagonal kernel checkpoints a pair: the
 * forward keeps the restart state of every K-th antidiagonal and the
 * backward recomputes K antidiagonals at a time, so lattice memory is about
 * (2 (Lx + Ly) / K + K + 3) (Lx + 1) (NBP + 1) doubles.  k = 0 (default):
 * K = max(4, ceil(sqrt(2 (Lx + Ly + 1)))) per pair, which minimises it;
 * k >= 4: K = max(4, min(k, Lx + Ly + 1)) for every pair (k in 1..3:
 * TKFMI_EINVAL).  Results do not depend on K.  The banded kernel's
 * recomputation is per band and takes no K.  Process-wide; query workspace
 * sizes after changing it.  The reference's counterpart is the segment
 * length of its rematerialisation, python/tkfmixdom/jax/train/vjp.py:627
 * (jax.checkpoint over the whole scan). */
int tkfmi_checkpoint_segment(int k);

/* Version of this header's ABI (TKFMI_ABI_VERSION). */
int tkfmi_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char* tkfmi_last_error(void);

/* Bytes of device workspace tkfmi_fb2d_batched needs for `prob`/`flags`. */
size_t tkfmi_fb2d_workspace_bytes(const tkfmi_problem* prob, uint32_t flags);

/*
 * Batched 2D pair-HMM forward/backward + expected transition counts.
 * One launch over the whole ragged batch.  Replaces, per pair:
 *   forward_backward_2d      dp/hmm.py:1834-1973 (logP, posteriors, n)
 *   forward_2d (logP only)   dp/hmm.py:1775-1801 (TKFMI_FORWARD_ONLY)
 * and the vmapped callers batched_fb_2d_per_chi (train_pfam.py:1674-1741)
 * and estep_batch_tkf92 (train/tkf92_svi_bw.py:263-334), without
 * geometric-bin padding (dp/hmm.py:66-91): every pair runs at its real
 * (Lx, Ly).  Emissions are looked up from per-group tables exactly as the
 * reference's builders define them (dp/hmm.py:628-851).
 *   n_trans[b] follows dp/hmm.py:1930-1971 (S,S); log_prob[b] is
 *   logsumexp_k F[Lx,Ly,k] + T[k,E] (dp/hmm.py:1911-1913).
 *   emit_counts[b] (TKFMI_WANT_EMIT_COUNTS) replaces the posterior
 *   reductions over the lattice: dom_match / dom_insert / dom_delete /
 *   agg_match (train_pfam.py:1705-1734; the reference's (K,A,A), (K,A)
 *   blocks are the [:20,:20] / [:20] parts of each group's 21-wide block),
 *   observation_counts_2d (train/likelihood.py:962-1028: I counts include
 *   row i=0, D counts column j=0) and the emission gradients
 *   (train/vjp.py:27-106); the MixDom2 class counts follow from them
 *   (train/likelihood.py:296-319).  Accumulated in LDS with FP64 atomics,
 *   so the summation order (not the value beyond rounding) may vary from
 *   run to run.
 * Requirements: state 0 has type S, some state has type E, at most 160
 * emitting (M/I/D) states, 0 <= Lx, Ly, Lx + Ly <= 32000.
 * Kernel choice (all compute the same quantities):
 *   - the TKF (S,M,I,D,E) layout and the MixDom layout (core/params.py:22,
 *     models/mixdom.py:69-77) with K*F <= 3, table emissions: the
 *     cell-per-lane kernels, by default as chains of workgroups per pair (a
 *     pair's 64-row bands spread over several compute units, one forward and
 *     one backward launch per wave-count class, workgroups ticketed in
 *     dependency order), or with TKFMI_CLASS_KERNEL one workgroup per pair;
 *   - other layouts with at most 15 emitting states, dense emissions, or
 *     TKFMI_GENERIC_KERNEL: the generic kernel;
 *   - 16..160 emitting states: MFMA transition and count contractions, by
 *     default as chains of workgroups per pair (a pair's 16-row bands on
 *     separate compute units, ticketed in dependency order; checkpointed
 *     bands recompute their lattice in the backward), or one workgroup per
 *     pair sweeping antidiagonals, which is what runs with
 *     TKFMI_CLASS_KERNEL and when checkpointing a batch of few bands
 *     (recomputation slots there would save less than half the lattice).
 * If a chained launch ever times out on an internal wait, the affected
 * pairs' log_prob is NaN and outputs.status (if given) is set.
 * Thread safety: calls from several host threads are safe as long as each
 * call has its own workspace and outputs; calls on different streams
 * overlap on the device (each caller stream gets its own library streams).
 */
int tkfmi_fb2d_batched(const tkfmi_problem* prob, const tkfmi_outputs* out,
                       void* workspace, size_t workspace_bytes,
                       void* stream /* hipStream_t */, uint32_t flags);

/* Bytes of device scratch behind tkfmi_fb2d_batched_noalloc (576 MiB): one
 * full forward/backward of a 2000 x 2000 pair at 17 states (518 MiB), a
 * minibatch of ~50 pairs of 250 x 250 at 17 states, or forward-only scoring
 * of pairs up to 2000 x 2000 at 162 states. */
#define TKFMI_NOALLOC_SCRATCH_BYTES ((size_t)576 << 20)

/*
 * tkfmi_fb2d_batched without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The scratch the kernels need comes from
 * a library-owned device array of TKFMI_NOALLOC_SCRATCH_BYTES, one per GPU,
 * held until the process ends; a batch needing more
 * (tkfmi_fb2d_workspace_bytes) is refused with TKFMI_EINVAL and a message.
 * Descriptors are copied from host memory on `stream` and every kernel runs
 * on `stream` (no library streams, events or pinned buffers are created).
 * TKFMI_TIME_KERNELS and TKFMI_REUSE_DESCRIPTORS are refused.  Results are
 * those of tkfmi_fb2d_batched.  All calls on one GPU share its scratch:
 * calls that overlap on one GPU (several streams or threads) race, and the
 * caller must serialise them.
 */
int tkfmi_fb2d_batched_noalloc(const tkfmi_problem* prob, const tkfmi_outputs* out,
                               void* stream /* hipStream_t */, uint32_t flags);

/* ------------------------------------------------------------------------
 * 1D alignment-constrained forward/backward (the MSA-guided DP).
 * ---------------------------------------------------------------------- */

/* A ragged batch of 1D problems: pair b has h_len[b] alignment columns,
 * rows h_off[b] .. h_off[b] + h_len[b] - 1 of the row-indexed inputs.
 * Emissions come either densely (log_emit != NULL: the reference's
 * log_emissions argument, NEG_INF = -1e30 where a state cannot emit) or
 * from the column types and characters plus per-group tables (log_emit ==
 * NULL: pair_hmm_emissions_constrained, dp/hmm.py:854-1100, computed on the
 * device from the same tables the 2D path uses). */
typedef struct tkfmi_problem1d {
  int32_t B;                   /* number of pairs                           */
  int32_t S;                   /* number of states, 1..256                  */
  int32_t s_idx;               /* init state (the reference's init_state)   */
  int32_t e_idx;               /* final state (final_state)                 */
  int32_t n_trans;             /* 1 (shared) or B (per pair)                */
  int32_t pad0;
  const double* log_trans;     /* device [n_trans,S,S] log transitions      */
  const double* log_emit;      /* device [rows,S] log emissions, or NULL    */
  const int32_t* state_types;  /* host [S] (constrained mode): 0=S 1=M 2=I 3=D 4=E */
  const int32_t* state_group;  /* host [S] (constrained mode)               */
  int32_t n_groups;            /* G (constrained mode)                      */
  int32_t n_tab;               /* 1 or B (constrained mode)                 */
  const double* log_tab;       /* device [n_tab,G,TKFMI_TAB_SZ] (constrained) */
  const uint8_t* col_type;     /* device [rows]: 1=M 2=I 3=D (constrained)  */
  const uint8_t* anc;          /* device [rows]: ancestor code at M/D columns */
  const uint8_t* desc;         /* device [rows]: descendant code at M/I columns */
  const int64_t* h_off;        /* host [B] first row of each pair           */
  const int32_t* h_len;        /* host [B] columns L of each pair           */
} tkfmi_problem1d;

/* Bytes of device workspace tkfmi_fb1d_batched needs for `prob`/`flags`. */
size_t tkfmi_fb1d_workspace_bytes(const tkfmi_problem1d* prob, uint32_t flags);

/*
 * Batched 1D forward/backward + expected transition counts.  Replaces, per
 * pair, forward_backward_1d_padded (dp/hmm.py:471-607, with forward_1d_scan
 * :404-437 and backward_1d_scan :440-468; TKFMI_FORWARD_ONLY gives the
 * forward-only fast path) and its vmapped caller batched_fb_per_chi
 * (train_pfam.py:1576-1610, used by _process_pairs_batched
 * train_pfam.py:2049-2603), at each pair's real length (no _pad_to_bin).
 *   log_prob[b]: logsumexp_k alpha[L-1,k] + T[k,final]; L = 0 gives
 *     T[init,final] and n[init,final] = 1 (dp/hmm.py:598-607).
 *   n_trans[b] (S,S): interior, terminal and initial terms of
 *     dp/hmm.py:549-596; the backward uses T with T[final,final] = 1
 *     (dp/hmm.py:532) as the reference does.
 *   posteriors (TKFMI_WANT_POSTERIORS): ragged [sum_b L_b, S], pair b at row
 *     sum_{b'<b} L_b', exp(alpha + beta - logP) (dp/hmm.py:547).
 *   emit_counts (TKFMI_WANT_EMIT_COUNTS, constrained mode only): [B,G,TAB_SZ]
 *     posterior mass per table entry, the input of the reference's
 *     per-domain / per-class count reductions (train_pfam.py:2338-2546).
 * Outputs n_trans / posteriors / emit_counts are accumulated (+=) and must
 * be zero-filled by the caller.  An impossible pair gets log_prob -1e30
 * and zero counts.
 */
int tkfmi_fb1d_batched(const tkfmi_problem1d* prob, const tkfmi_outputs* out,
                       void* workspace, size_t workspace_bytes,
                       void* stream /* hipStream_t */, uint32_t flags);

/* ------------------------------------------------------------------------
 * Null-state count restoration, sufficient statistics and BDI statistics.
 * ---------------------------------------------------------------------- */

/* Shared MixDom structural parameters (host arrays). */
typedef struct tkfmi_mixdom_model {
  int32_t K, F;             /* domains (1..16), fragments per domain (>= 1)   */
  double main_ins, main_del;
  const double* dom_ins;    /* host [K]                                      */
  const double* dom_del;    /* host [K]                                      */
  const double* dom_w;      /* host [K]                                      */
  const double* frag_w;     /* host [K,F]                                    */
  const double* ext;        /* host [K,F,F]; MixDom1 (K,F) rates go on the
                               diagonal (models/mixdom.py:284-287)            */
} tkfmi_mixdom_model;

/* Doubles per pair of the statistics vector written by
 * tkfmi_restore_counts_batched, in this order:
 *   top_5x5 [5,5] | dom_M_5x5 [K,5,5] | dom_kappa [K] | dom_1mkappa [K] |
 *   ext [K,F,F] | term [K,F] | dom_w [K] | frag_w [K,F]
 * (= 25 + 28K + K F (F + 2)). */
int32_t tkfmi_suffstats_size(int32_t K, int32_t F);

/* Bytes of device workspace tkfmi_restore_counts_batched needs. */
size_t tkfmi_restore_workspace_bytes(const tkfmi_mixdom_model* m, int32_t B);

/*
 * Per pair b: rebuild the fully exploded MixDom machine at t[b]
 * (models/exact_suffstats_batch.py:61-253, N = 8 + 8K + 15KF states),
 * eliminate its null groups in chain order (:468-482, :272-282), restore the
 * collapsed expected counts n_chi[b] to the exploded machine by ghost usage
 * (:285-339) and extract the per-group statistics (:347-458).  Replaces
 * exact_suffstats_batch (:485-541) and the per-pair loop of
 * exact_suffstats_per_pair_batch (models/exact_suffstats.py:266-386).
 *   t       device [B]
 *   n_chi   device [B, N_c, N_c], N_c = 2 + 5KF (collapsed chi layout, as
 *           written by tkfmi_fb2d_batched)
 *   stats   device [B, tkfmi_suffstats_size(K, F)]
 *   bdi     device [B, 1 + K, 4] or NULL: per level (top, then each domain)
 *           E[B], E[D], E[S] (core/bdi.py:682-778) and the observation time
 *           T (top: t * sum top_5x5[:,E]; domain d: t * (sum dom_M_5x5[d][S,:]
 *           + dom_1mkappa[d]), models/exact_suffstats.py:353-381).
 * FP64 throughout, one workgroup per pair, stream-ordered.
 */
int tkfmi_restore_counts_batched(const tkfmi_mixdom_model* m, int32_t B, const double* t,
                                 const double* n_chi, double* stats, double* bdi,
                                 void* workspace, size_t workspace_bytes, void* stream);

/*
 * BDI statistics (E[B], E[D], E[S]) of B TKF91-semantics 5x5 count
 * matrices at shared (lam, mu): tkf91_stats_from_counts_batch
 * (core/bdi.py:800-888), general or equal-rate (|1 - lam/mu| < 1e-10)
 * branch.  n5 device [B,5,5], t and T device [B], out device [B,3].
 */
int tkfmi_bdi_stats_batched(int32_t B, double lam, double mu, const double* n5, const double* t,
                            const double* T, double* out, void* stream);

/* Doubles per pair written by tkfmi_tkf92_estep_stats. */
#define TKFMI_TKF92_STATS 35

/*
 * TKF92 E-step statistics of B pairs from their chi counts n_chi [B,5,5]
 * (as written by tkfmi_fb2d_batched for the TKF layout) at per-pair t:
 * the self-loop split _tkf92_chi_resolve_core (train/tkf92_svi_bw.py:75-123),
 * the log-kappa coefficients (:160-167) and the clamped BDI statistics of
 * estep_batch_tkf92 (:263-334).  out device [B, TKFMI_TKF92_STATS]:
 *   resolved [5,5] | ext_count | notext_count | L | M | E_B | E_D | E_S |
 *   T (= t) | pad [2]
 */
int tkfmi_tkf92_estep_stats(int32_t B, double lam, double mu, double ext, const double* t,
                            const double* n_chi, double* out, void* stream);

/*
 * Holmes-Rubin weighted statistics per (pair, class): expected dwell times
 * W [A] and substitution counts U [A,A] of the CTMC paths behind the pair's
 * expected match counts (core/ctmc.py:173-326 holmes_rubin_weighted_stats;
 * the per-pair-per-domain batch of train_pfam.py:1991-2046).  The caller
 * supplies, per class c, the eigendecomposition S_c = V diag(l) V^T of the
 * symmetrised rate matrix S_c = D^1/2 Q_c D^-1/2 (D = diag(pi_c), made
 * exactly symmetric as (S + S^T)/2), shared by all pairs.
 *   eigvec device [C,A,A] (row a, column k = component a of eigenvector k)
 *   eigval device [C,A], pi device [C,A], Q device [C,A,A], t device [B],
 *   mc device [B,C,A,A] (match counts), W device [B,C,A], U device [B,C,A,A].
 * A <= 32.
 */
int tkfmi_hr_stats_batched(int32_t B, int32_t C, int32_t A, const double* eigvec,
                           const double* eigval, const double* pi, const double* Q, const double* t,
                           const double* mc, double* W, double* U, void* stream);

/* ------------------------------------------------------------------------
 * Per-pair transition matrices and emission tables on the device.
 * ---------------------------------------------------------------------- */
#define TKFMI_MODEL_TKF91 0
#define TKFMI_MODEL_TKF91_COND 1
#define TKFMI_MODEL_TKF92 2
#define TKFMI_MODEL_MIXDOM 3
#define TKFMI_EMIT_SHARED 0      /* log(x + 1e-30), one table (dp/hmm.py:628-670)      */
#define TKFMI_EMIT_PER_DOMAIN 1  /* log(max(x, 1e-30)), group = domain (:786-851)      */
#define TKFMI_EMIT_PER_CLASS 2   /* class mixture per (domain, fragment) (:673-783)    */

typedef struct tkfmi_builder {
  int32_t model;      /* TKFMI_MODEL_*                                       */
  int32_t emission;   /* TKFMI_EMIT_*                                        */
  int32_t log_floor;  /* transitions: 0 = safe_log (NEG_INF for 0), 1 = log(max(x, 1e-30)) */
  int32_t C;          /* substitution components: 1 shared, K per-domain, classes per-class */
  int32_t A;          /* alphabet (<= 20)                                    */
  int32_t pad;
  double lam, mu, ext;                 /* TKF models                          */
  const tkfmi_mixdom_model* mix;       /* host, TKFMI_MODEL_MIXDOM            */
  const double* Q;       /* device [C,A,A] rate matrices (rows sum to zero;
                            reversibility is not assumed)                    */
  const double* pi;      /* device [C,A] equilibria (emission pi factors)    */
  const double* class_dist;  /* device [K,F,C] (per-class) or NULL           */
} tkfmi_builder;

/* Bytes of device workspace tkfmi_build_params_batched needs. */
size_t tkfmi_build_workspace_bytes(const tkfmi_builder* b, int32_t B);

/*
 * For every pair b at t[b]: log_trans [B,S,S] (TKF91 / TKF91-cond / TKF92,
 * core/params.py:29-108, or the collapsed MixDom chi, models/mixdom.py:252-423,
 * obtained as the reduced matrix of the exploded machine) and log_tab
 * [B,G,TKFMI_TAB_SZ] built from P_c(t) = expm(Q_c t) (core/ctmc.py:31-60),
 * evaluated on the device by the Pade scaling-and-squaring rule of
 * jax.scipy.linalg.expm (degree 3..13 by the 1-norm of Q t, floor(log2(
 * norm / 5.3719...)) squarings, NaN beyond 16 squarings) for any Q.
 * Replaces the per-pair builders of train_pfam.py:1613-1664, 1940-1988 and
 * train/tkf92_svi_bw.py:148-152.  G = 1 (shared), K (per-domain) or K*F
 * (per-class).  One workgroup per pair.
 */
int tkfmi_build_params_batched(const tkfmi_builder* b, int32_t B, const double* t,
                               double* log_trans, double* log_tab, void* workspace,
                               size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Builders of the gradient path with their adjoints (tkfmi/vjp.py's
 * differentiable log-likelihoods, the reference's train/vjp.py surfaces).
 * The reference differentiates its JAX builders by autodiff; these give the
 * same derivatives from a few kernels per step.  All arrays are device
 * arrays, row-major, FP64; calls are asynchronous on `stream`.
 * ---------------------------------------------------------------------- */

/*
 * For q < n: T[q] (5 x 5, states S M I D E) = TKF91 (mode 0), the
 * ancestor-length-conditioned TKF91 (mode 1) or TKF92 (mode 2) at
 * (lam[q], mu[q], t[q], ext[q]) (core/params.py:29-108 with core/bdi.py:40-97:
 * the same guard constants and the |1 - lambda/mu| < 1e-4 limits), and
 * J[q] (5 x 5 x 4) = dT / d(lam, mu, t, ext).  ext may be NULL unless
 * mode = 2.  Replaces autodiff through the builders of train/vjp.py:110-329.
 */
int tkfmi_tkf_trans_jac(const double* lam, const double* mu, const double* t, const double* ext,
                        int64_t n, int32_t mode, double* T, double* J, void* stream);

/* Bytes of device workspace tkfmi_chi_forward / tkfmi_chi_backward need. */
size_t tkfmi_chi_workspace_bytes(int64_t B, int32_t nk, int32_t nz);

/*
 * Collapsed MixDom chi of B pairs (models/mixdom.py:252-423, as the reduced
 * matrix of the exploded machine, models/exact_suffstats_batch.py:61-282):
 * the exploded matrix T (N0 x N0) has entry flat[e] = r N0 + c equal to
 * atoms[b, a1[e]] * atoms[b, a2[e]] for e < n_e (no flat index twice) and
 * 0 elsewhere; kmap[s] / zmap[s] give state s's index among the nk kept /
 * nz eliminated states (-1 otherwise).  chi [B,nk,nk] = T_KK + T_KZ W with
 * W [B,nz,nk] = (I - T_ZZ)^-1 T_ZK (kept for the backward).  I - T_ZZ must
 * be a nonsingular M-matrix (it is for every MixDom machine: the null states
 * reach End), which lets the solve run without pivoting.
 */
int tkfmi_chi_forward(const double* atoms, int64_t B, int32_t n_atoms, const int32_t* a1,
                      const int32_t* a2, const int32_t* flat, int32_t n_e, const int32_t* kmap,
                      const int32_t* zmap, int32_t N0, int32_t nk, int32_t nz, double* chi,
                      double* W, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Adjoint of tkfmi_chi_forward: gatoms [B,n_atoms] = d loss / d atoms from
 * G [B,nk,nk] = d loss / d chi and the forward's W.  nz + n_atoms <= 8192.
 */
int tkfmi_chi_backward(const double* atoms, int64_t B, int32_t n_atoms, const int32_t* a1,
                       const int32_t* a2, const int32_t* flat, int32_t n_e, const int32_t* kmap,
                       const int32_t* zmap, int32_t N0, int32_t nk, int32_t nz, const double* G,
                       const double* W, double* gatoms, void* workspace, size_t workspace_bytes,
                       void* stream);

#define TKFMI_EXPM_MAX_N 40

/*
 * out[q] = exp(A[q]) for count n x n matrices, n <= TKFMI_EXPM_MAX_N, by the
 * rule of jax.scipy.linalg.expm (the reference's transition_matrix,
 * core/ctmc.py:31-60; see tkfmi_build_params_batched).  One workgroup per
 * matrix.
 */
int tkfmi_expm_batched(const double* A, int64_t count, int32_t n, double* out, void* stream);

/*
 * Adjoint of exp: out[q] = d loss / d A[q] for G[q] = d loss / d exp(A[q]),
 * i.e. the Frechet derivative of exp at A[q]^T in the direction G[q] (the
 * upper-right block of exp([[A^T, G], [0, A^T]])), computed with the
 * degree-13 Pade approximant and the squarings exp(A[q]) takes;
 * 2 n <= TKFMI_EXPM_MAX_N.
 */
int tkfmi_expm_vjp_batched(const double* A, const double* G, int64_t count, int32_t n,
                           double* out, void* stream);

/* ------------------------------------------------------------------------
 * Precompiled X/A/Y pairs decoded on the device.
 * ---------------------------------------------------------------------- */

/* status bits of a decoded pair (0 = valid, util/pair_format.py:193-223) */
#define TKFMI_PAIR_BAD_X 1      /* a character of X outside "ACDEFGHIKLMNPQRSTVWY" */
#define TKFMI_PAIR_BAD_Y 2      /* a character of Y outside the 20 amino acids     */
#define TKFMI_PAIR_BAD_LEN_X 4  /* the M and D run lengths do not sum to |X|       */
#define TKFMI_PAIR_BAD_LEN_Y 8  /* the M and I run lengths do not sum to |Y|       */
#define TKFMI_PAIR_BAD_RUN 16   /* a run length above 2^24                         */

/* The text of B records (util/pair_format.py:93-137: 'x', 'a', 'y'), as
 * uploaded: the X strings, the Y strings and the run-length alignment
 * strings ("M5I1M3D2") each concatenated into one device byte buffer. */
typedef struct tkfmi_pair_text {
  int32_t B;
  int32_t pad;
  const uint8_t* x_text;    /* device ASCII                                   */
  const uint8_t* y_text;    /* device ASCII                                   */
  const uint8_t* a_text;    /* device ASCII                                   */
  const int64_t* h_x_off;   /* host [B]: pair b's X at x_text + h_x_off[b]     */
  const int32_t* h_x_len;   /* host [B]: |X| = Lx                              */
  const int64_t* h_y_off;   /* host [B]                                        */
  const int32_t* h_y_len;   /* host [B]: |Y| = Ly                              */
  const int64_t* h_a_off;   /* host [B]                                        */
  const int32_t* h_a_len;   /* host [B]: bytes of the alignment string         */
} tkfmi_pair_text;

/* Decoded pairs.  x_codes / y_codes have the offsets of x_text / y_text and
 * are the x_cat / y_cat of a tkfmi_problem.  Pair b's alignment columns start
 * at sum_{b'<b} (Lx + Ly) (its capacity; it has n_cols[b] <= Lx + Ly) and are
 * the col_type / anc / desc of a tkfmi_problem1d. */
typedef struct tkfmi_pair_columns {
  uint8_t* x_codes;   /* device: codes 0..19 (20 for an invalid character)    */
  uint8_t* y_codes;   /* device                                               */
  uint8_t* col_type;  /* device or NULL (then no columns): 1=M 2=I 3=D          */
  uint8_t* anc;       /* device or NULL: ancestor code at M/D columns, else 0   */
  uint8_t* desc;      /* device or NULL: descendant code at M/I columns, else 0 */
  int32_t* n_cols;    /* device [B]: alignment columns (0 if the lengths fail)  */
  int32_t* status;    /* device [B]: TKFMI_PAIR_* bits, 0 for a valid pair      */
} tkfmi_pair_columns;

/* Bytes of device workspace tkfmi_pairs_decode needs. */
size_t tkfmi_pairs_decode_workspace_bytes(const tkfmi_pair_text* in);

/*
 * Decodes B precompiled pairs in one launch, one workgroup per pair:
 * decode_pair (util/pair_format.py:140-190) with rle_decode (:69-90; a run
 * is an M, I or D followed by digits, other text is skipped as the
 * reference's regex skips it) and the checks of validate_record (:193-223),
 * for the pairs PrecompiledPairSource streams (train_pfam.py:984-1075).
 * Stream-ordered; status and n_cols are written for every pair.
 */
int tkfmi_pairs_decode(const tkfmi_pair_text* in, const tkfmi_pair_columns* out, void* workspace,
                       size_t workspace_bytes, void* stream /* hipStream_t */);

/* One kernel launch of a timed call. */
typedef struct tkfmi_kernel_time {
  char name[64];    /* kernel, e.g. "fbv2_kernel<LMix<3,1>,8>"            */
  int32_t n_pairs;  /* pairs processed by the launch                       */
  int32_t pad;
  int64_t cells;    /* sum over those pairs of (Lx+1)(Ly+1)                */
  double ms;        /* HIP-event duration on the stream it ran on          */
} tkfmi_kernel_time;

/*
 * Measurement aid with no counterpart in the reference.  Durations of the
 * kernels launched by the last tkfmi_fb2d_batched call on this thread that
 * had TKFMI_TIME_KERNELS set, measured with HIP events recorded on the
 * stream each kernel was launched on (the cell-per-lane path launches one
 * kernel per wave-count class on library-owned streams).  Waits for those
 * kernels.  Returns the number of entries written (at most `max`) or a
 * negative TKFMI_E* code.  Not meant for concurrent timed calls.
 */
int tkfmi_last_kernel_times(tkfmi_kernel_time* out, int32_t max);

/*
 * Tuning of the chained 2D kernels; no counterpart in the reference (its
 * 2D path is one vmapped JAX call per padded shape).  Process-wide, applies
 * to later calls (a call's workspace size depends on it: query again after
 * changing it).  A pair with at least long_w bands whose cells are at least
 * long_frac x the batch's largest pair runs as a chain of long_w-wave
 * workgroups on its own stream; the other pairs with at least main_w bands
 * as chains of main_w-wave workgroups; shorter pairs with 2- and 1-wave
 * workgroups.  long_one_launch != 0 runs that long class on the one-launch
 * schedule (see tkfmi_chain_schedule) whatever the call's size, the other
 * classes as that schedule decides.  long_w in 1..4, main_w in 2..4,
 * long_frac >= 0 (0: no long class).  Defaults 3, 3, 0.7, 1 (the few
 * longest pairs of a batch on the one-launch schedule).  tkfmi.autotune
 * chooses them by timing the caller's batch.  Returns TKFMI_OK or
 * TKFMI_EINVAL.
 */
int tkfmi_chain_config(int32_t long_w, int32_t main_w, double long_frac, int32_t long_one_launch);

/*
 * Schedule of the chained 2D kernels; no counterpart in the reference (it
 * runs a pair's forward, then its backward: dp/hmm.py:1834-1973).  In the
 * one-launch schedule the forward and the backward of every pair run at the
 * same time, from opposite corners, and meet at antidiagonal
 * C = round(cut (Lx + Ly)): the backward takes the expected counts of the
 * cells i + j <= C, the forward those of the cells past C, both scaled by
 * the P(x, y) of the transitions crossing the cut.  A pair then takes about
 * max(C, Lx + Ly - C) steps of each sweep instead of Lx + Ly of both in
 * turn, but all its workgroups stay resident together, so the schedule is
 * used only when the call's workgroups (two per band group of each pair)
 * number at most max_workgroups (-1: twice the device's compute units, the
 * number it holds at once); larger calls, calls asking for posteriors or
 * for log P only run the forward launch, then the backward launch.  cut = 0
 * never uses it.  Process-wide, applies to later calls (query the workspace
 * size again: the one-launch schedule adds 256 (Lx + 1) bytes per pair and
 * a second set of boundary rows).  Defaults 0.45 and -1, or TKFMI_CHAIN_CUT
 * and TKFMI_CHAIN_BIDI_MAX_WG.  Returns TKFMI_OK or TKFMI_EINVAL (cut
 * outside [0, 1]).
 */
int tkfmi_chain_schedule(double cut, int32_t max_workgroups);

/* Result of tkfmi_autotune: the candidates timed, in order (the first is
 * the configuration in force before the call), their ms per call (-1: not
 * run because its workspace exceeded the one given) and the chosen one. */
typedef struct tkfmi_tune_result {
  int32_t layout;      /* internal layout id of the key, -1 if not tunable     */
  int32_t length_bin;  /* longest sequence < 128, 256, 512, 1024, 2048, more   */
  int32_t n;           /* candidates timed (0: one configuration only)         */
  int32_t best;        /* index of the chosen candidate                        */
  int32_t long_w[16];
  int32_t main_w[16];
  double long_frac[16];
  double ms[16];
  int32_t long_one[16]; /* the long class on the one-launch schedule         */
} tkfmi_tune_result;

/* Workspace tkfmi_autotune needs for `prob` (the largest over candidates). */
size_t tkfmi_autotune_workspace_bytes(const tkfmi_problem* prob, uint32_t flags);

/*
 * Autotuner of the chained 2D kernels; no counterpart in the reference
 * (its 2D path is one vmapped JAX call per padded shape).  Times the
 * candidate class configurations (waves per workgroup of the long and main
 * classes, the long-class cell fraction) on the caller's own batch -- an
 * untimed warm-up, then two round-robin passes of `reps` timed calls per
 * candidate (each keeps its faster pass), HIP events on `stream`,
 * synchronous -- and keeps the fastest (the configuration already in force unless
 * another is more than 3 % faster) for the batch's key: its state layout
 * and the length bin of its longest sequence.  Later calls whose batch has
 * that key use it (query their workspace size after tuning: it depends on
 * the configuration); other keys keep tkfmi_chain_config's setting (which
 * clears every tuned key).  `out` receives the batch's results.  Layouts
 * without a tunable configuration (the generic and 16..160-state kernels,
 * dense emissions) return TKFMI_OK with res->n = 0.
 */
int tkfmi_autotune(const tkfmi_problem* prob, const tkfmi_outputs* out, void* workspace,
                   size_t workspace_bytes, void* stream /* hipStream_t */, uint32_t flags,
                   int32_t reps, tkfmi_tune_result* res);

#ifdef __cplusplus
}
#endif

#endif /* TKFMI_H */
