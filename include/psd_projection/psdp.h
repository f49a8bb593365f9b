/*
 * psdp.h -- C ABI of libpsdp, the MI355X (gfx950) PSD-projection library.
 *
 * Every entry point is `extern "C"`, takes plain device pointers and sizes,
 * and returns a psdp_status_t.  Matrices are dense, symmetric, column-major
 * n x n with leading dimension n, resident in device memory.  Work is
 * enqueued on the handle's HIP stream.
 *
 * Drop-in map (reference = ComputationalRobotics/psd_projection,
 * the headers in include/psd_projection; cuBLAS/cuSOLVER handles become one
 * psdp_handle_t; `float* workspace` becomes (ws, ws_bytes)):
 *
 *   psdp_composite_fp32                  <- composite_FP32              composite_FP32.h:12-17
 *   psdp_composite_fp32_auto_scale       <- composite_FP32_auto_scale   composite_FP32.h:24-30
 *   psdp_composite_fp16                  <- composite_FP16              composite_FP16.h:12-18
 *   psdp_composite_fp16_auto_scale       <- composite_FP16_auto_scale   composite_FP16.h:26-33
 *   psdp_composite_fp32_emulated         <- composite_FP32_emulated     composite_FP32_emulated.h:17-22
 *   psdp_composite_fp32_emulated_auto_scale
 *                                        <- composite_FP32_emulated_auto_scale
 *                                                                       composite_FP32_emulated.h:29-35
 *   psdp_approximate_two_norm            <- approximate_two_norm        lanczos.h:11-17
 *   psdp_eig_fp64_psd                    <- eig_FP64_psd                eig_FP64_psd.h:14-20
 *   psdp_eig_fp32_psd                    <- eig_FP32_psd                eig_FP32_psd.h:14-20
 *   psdp_compute_eigenpairs              <- compute_eigenpairs          lanczos.h:33-43
 *   psdp_lobpcg                          <- lobpcg                      lobpcg.h:16-28
 *   psdp_composite_fp32_auto_scale_deflate
 *                                        <- composite_FP32_auto_scale_deflate
 *                                                                       composite_FP32.h:36-45
 *   psdp_composite_fp16_auto_scale_deflate
 *                                        <- composite_FP16_auto_scale_deflate
 *                                                                       composite_FP16.h:39-50
 *   psdp_eig_fp64_deflate                <- eig_FP64_deflate            eig_FP64_psd.h:22-31
 *
 * Additions with no reference counterpart: float32 in/out projections with a
 * per-step precision schedule (psdp_project_f32/_f64), grouped projection of
 * many blocks (psdp_project_batched_f32), workspace sizing and the handle.
 *
 * Error convention: the reference prints CHECK_* failures and continues
 * (include/psd_projection/check.h:18-57); here every call returns a status and
 * psdp_last_error() describes the last failure of the calling thread.
 */
#ifndef PSD_PROJECTION_PSDP_H
#define PSD_PROJECTION_PSDP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum psdp_status {
  PSDP_OK = 0,
  PSDP_ERR_INVALID_ARG = 1,
  PSDP_ERR_NOMEM = 2,
  PSDP_ERR_NONFINITE = 3,
  PSDP_ERR_UNSUPPORTED = 4,
  PSDP_ERR_DEVICE = 5
} psdp_status_t;

/* Operand format of the products of one filter step. */
typedef enum psdp_format {
  PSDP_FMT_F32 = 0,    /* float32 operands on the FP32 MFMA                       */
  PSDP_FMT_F16 = 1,    /* binary16 operands, float32 accumulation                  */
  PSDP_FMT_F16X3 = 2,  /* 2^8 v = hi + lo binary16 split, 3 products (~22-bit)     */
  PSDP_FMT_BF16 = 3,   /* bfloat16 operands                                        */
  PSDP_FMT_BF16X3 = 4, /* hi + lo bfloat16 split, 3 products (~16-bit)             */
  PSDP_FMT_BF16X9 = 5, /* hi + mid + lo bfloat16 split, 9 products (~fp32)         */
  PSDP_FMT_E4M3 = 6,   /* FP8 e4m3 operand values, per-tensor scale 2^8 (|a| <= 1  */
                       /* after the bound), products on the binary16 MFMA          */
  PSDP_FMT_F16MX = 7   /* hi binary16 of 2^8 v for hi.hi, e4m3 copies of hi and lo */
                       /* with one power-of-two scale per 32 values (MX blocks)    */
                       /* for the cross terms on the scaled FP8 MFMA: 2 binary16-  */
                       /* MFMA equivalents per k instead of FP16x3's 3 (~15-bit)   */
} psdp_format_t;

/* Coefficient table of the composite polynomial chain. */
typedef enum psdp_table {
  PSDP_TABLE_FP32 = 0, /* 10 steps, src/composite_FP32.cu:73-86 (scale 1.001, steps 0-7) */
  PSDP_TABLE_FP16 = 1  /* 7 steps,  src/composite_FP16.cu:93-103 (scale 1.01, all steps) */
} psdp_table_t;

/* Named methods (the reference's MEX method strings plus the shipped
 * mixed-precision schedule). */
typedef enum psdp_method {
  PSDP_METHOD_COMPOSITE_FP32 = 0,          /* FP32 table, FP32 MFMA everywhere          */
  PSDP_METHOD_COMPOSITE_FP16 = 1,          /* FP16 table, FP16 operands everywhere      */
  PSDP_METHOD_COMPOSITE_FP32_EMULATED = 2, /* FP32 table, BF16x9 everywhere             */
  PSDP_METHOD_MIXED = 3,                   /* FP16 table, FP16x3 on steps 0-3, FP16 rest */
  PSDP_METHOD_CUSTOM = 4,                  /* table + formats taken from psdp_config_t  */
  PSDP_METHOD_MIXED_MX = 5                 /* FP16 table, F16MX on steps 0-3, FP16 rest  */
} psdp_method_t;

#define PSDP_MAX_STEPS 16

/* Residual of the Lanczos bound up = sqrt(theta + ||r||) (theta: largest
 * Ritz value of A^T A, y: its unit Ritz vector). */
typedef enum psdp_bound_mode {
  PSDP_BOUND_REFERENCE = 0, /* r = A y - theta y, as src/lanczos.cu:255 computes it (default) */
  PSDP_BOUND_TIGHT = 1      /* r = A (A y) - theta y: the Ritz pair's own residual (one more   */
                            /* matrix-vector product); up then exceeds ||A||_2 by the Ritz     */
                            /* error only, so clustered spectra keep more of [-1, 1]            */
} psdp_bound_mode_t;

typedef struct psdp_config {
  int method;                     /* psdp_method_t                                   */
  int table;                      /* psdp_table_t (CUSTOM only)                      */
  int step_format[PSDP_MAX_STEPS];/* psdp_format_t per step (CUSTOM only)            */
  int final_format;               /* psdp_format_t of W = X S (CUSTOM only)          */
  int auto_scale;                 /* 1: Lanczos bound + rescale (the *_auto_scale)   */
  int lanczos_iters;              /* default 20 (lanczos.h:16), at most 64           */
  double lanczos_tol;             /* default 1e-10 (lanczos.h:16)                    */
  double scale_margin;            /* s = up * margin; default 1.0 (reference)        */
  int synchronize;                /* 1: wait for the stream and check device flags   */
  int average_products;           /* CUSTOM: 1-piece 16-bit steps use (C + C^T)/2    */
  int bound_mode;                 /* psdp_bound_mode_t (default REFERENCE)           */
} psdp_config_t;

typedef struct psdp_handle_st* psdp_handle_t;

/* Handle: device, stream (default: a new non-blocking stream owned by the
 * handle), workspace arena, Lanczos state, cached tile lists.
 * psdp_set_stream: later calls enqueue on hip_stream, which first waits (on
 * the device) for the work already enqueued through the handle, because that
 * work and the next call share the handle's workspace.  One handle serves one
 * host thread at a time; use one handle per concurrent stream of work.
 * Calls with cfg->synchronize != 0 (the default, and every reference drop-in)
 * return after the result is written; with synchronize == 0 they return once
 * the work is enqueued (psdp_synchronize waits for it). */
psdp_status_t psdp_create(psdp_handle_t* handle, int device);
psdp_status_t psdp_destroy(psdp_handle_t handle);
psdp_status_t psdp_set_stream(psdp_handle_t handle, void* hip_stream);
psdp_status_t psdp_get_stream(psdp_handle_t handle, void** hip_stream);
psdp_status_t psdp_set_seed(psdp_handle_t handle, uint64_t seed);
psdp_status_t psdp_synchronize(psdp_handle_t handle);
const char* psdp_last_error(void);
const char* psdp_status_string(psdp_status_t status);
psdp_status_t psdp_default_config(psdp_config_t* cfg, int method);

/* Bytes of workspace a projection of size n with cfg needs (ws may then be
 * passed explicitly; ws == NULL uses the handle's arena). */
psdp_status_t psdp_workspace_size(const psdp_config_t* cfg, int64_t n, size_t* bytes);

/* Lanczos bound of the last auto-scaled projection (waits for the stream). */
psdp_status_t psdp_last_bound(psdp_handle_t handle, double* lo, double* up, int* iterations);
/* Per-block bounds of the handle's last batched (or packed) call: lo, up
 * (host arrays of count doubles, either may be NULL) and iterations (host,
 * count ints, may be NULL); count must equal that call's block count.  Waits
 * for the stream.  Valid until the next call that uses the handle's arena. */
psdp_status_t psdp_last_bounds_batched(psdp_handle_t handle, double* lo, double* up, int* iterations, int count);

/* ---- drop-ins of the reference entry points (in place, FP64) ----
 * Like the reference, these return after the result is written: they wait
 * for the handle's stream and report a non-finite input, bound or result
 * (an unscaled spectrum outside [-1, 1] makes the chain diverge,
 * include/psd_projection/composite_FP32.h:11) as PSDP_ERR_NONFINITE. */
psdp_status_t psdp_composite_fp32(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes);
psdp_status_t psdp_composite_fp32_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                             size_t ws_bytes);
psdp_status_t psdp_composite_fp16(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes);
psdp_status_t psdp_composite_fp16_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                             size_t ws_bytes);
psdp_status_t psdp_composite_fp32_emulated(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                           size_t ws_bytes);
psdp_status_t psdp_composite_fp32_emulated_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                                      size_t ws_bytes);
/* lo/up are host pointers, as in the reference (the call blocks). */
psdp_status_t psdp_approximate_two_norm(psdp_handle_t h, const double* A, int64_t n, double* lo, double* up,
                                        int64_t max_iter, double tol);

/* ---- general projections: Y = Pi(X); X == Y allowed; column-major, ld = n ---- */
psdp_status_t psdp_project_f64(psdp_handle_t h, const double* X, double* Y, int64_t n,
                               const psdp_config_t* cfg, void* ws, size_t ws_bytes);
psdp_status_t psdp_project_f32(psdp_handle_t h, const float* X, float* Y, int64_t n,
                               const psdp_config_t* cfg, void* ws, size_t ws_bytes);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory inside a call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): psdp_project_f32 with its workspace,
 * Lanczos state and tile lists taken from a fixed 64 MiB device array that the
 * library holds on every GPU it runs on until the process ends; the call
 * makes no hipMalloc/hipFree.  Sizes whose workspace does not fit are
 * rejected with PSDP_ERR_INVALID_ARG (n up to 1664 for the mixed schedule,
 * 1920 for the FP32 and FP16 methods, 1408 for the emulated one).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (any
 * handle, any stream) race. */
psdp_status_t psdp_project_f32_noalloc(psdp_handle_t h, const float* X, float* Y, int64_t n,
                                       const psdp_config_t* cfg);

/* One symmetric product of X, Y (symmetric n x n float32, column-major) with
 * operands rounded to `format`: C = upper triangle of X Y mirrored, or with
 * average != 0 (binary16/bfloat16 only) C = (X Y + Y X)/2.  Exposes the
 * filter-chain GEMM kernel on its own. */
psdp_status_t psdp_sym_product_f32(psdp_handle_t h, const float* X, const float* Y, float* C, int64_t n,
                                   int format, int average);

/* ---- profiling: HIP events around every launch of the handle, by kind ---- */
typedef enum psdp_prof_kind {
  PSDP_PROF_PROD_F32 = 0, /* + psdp_format_t: symmetric products by operand format */
  PSDP_PROF_LANCZOS = 8,  /* the whole spectral bound                                */
  PSDP_PROF_PREP = 9,     /* symmetrize/pad + scale/split                            */
  PSDP_PROF_EIGENPAIRS = 10, /* psdp_compute_eigenpairs (whole call)                */
  PSDP_PROF_LOBPCG = 11,  /* psdp_lobpcg (whole call)                                */
  PSDP_PROF_RANK_UPDATE = 12, /* X += V diag(c) V^T of the deflated projections      */
  PSDP_PROF_GRAPH = 13,   /* a replayed batched graph (whole device sequence)       */
  PSDP_PROF_KINDS = 14
} psdp_prof_kind_t;
/* on != 0 starts recording (and clears old records); 0 stops. */
psdp_status_t psdp_profile_enable(psdp_handle_t h, int on);
/* Waits for the stream; returns launches and summed milliseconds of a kind. */
psdp_status_t psdp_profile_summary(psdp_handle_t h, int kind, int* launches, double* total_ms);

/* ---- batched block-diagonal cone: every block projected independently ---- */
typedef struct psdp_block {
  const float* X; /* device, column-major n x n */
  float* Y;       /* device output (may equal X) */
  int64_t n;
} psdp_block_t;
psdp_status_t psdp_project_batched_f32(psdp_handle_t h, const psdp_block_t* blocks, int count,
                                       const psdp_config_t* cfg);
/* Packed block-diagonal cone, the on-device form an ADMM/SDP solver keeps:
 * count column-major blocks of sizes[b] x sizes[b] floats stored back to back
 * (block b starts at element sum_{c<b} sizes[c]^2) in one device buffer X;
 * the projected blocks go to Y in the same layout (Y == X projects in place).
 * Same arithmetic as psdp_project_batched_f32 on those blocks.  No reference
 * counterpart: the reference projects one matrix per call (its intended
 * caller, an SDP solver, include/psd_projection/check.h:10, loops over cones). */
psdp_status_t psdp_project_packed_f32(psdp_handle_t h, const float* X, float* Y, const int64_t* sizes, int count,
                                      const psdp_config_t* cfg);
/* ---- continuous batching: a session of blocks that join and leave between products ----
 * No reference counterpart: the reference projects one matrix per call; its
 * intended caller, an ADMM/SDP solver (include/psd_projection/check.h:10),
 * projects many cone blocks every iteration.
 * A session keeps a changing set of float32 blocks in flight on the handle's
 * stream.  psdp_session_add admits a block (X, Y: device, column-major n x n;
 * Y is written when the block completes) and returns its ticket; its
 * workspace comes from the session's arena, sized once at creation, so
 * nothing is allocated per block: PSDP_ERR_NOMEM says the arena has no room
 * for it now (retry after completions), PSDP_ERR_INVALID_ARG that it never
 * will.  Each psdp_session_step enqueues the prologue (symmetrize, Lanczos
 * bound, scale/split) of the blocks added since the last step on the
 * session's second stream (at most four prologues in flight), and one tick on
 * the handle's stream: ONE filter-chain product for every block whose
 * prologue has completed, one grouped launch per operand format and tile
 * class.  A block thus starts at the next
 * product boundary instead of waiting for the blocks ahead of it, and leaves
 * after its output product; *in_flight_ticks (may be NULL) counts the ticks
 * enqueued and not yet complete.  psdp_session_poll returns (up to max)
 * tickets whose outputs are complete, *active (may be NULL): the blocks not
 * yet complete, and *in_flight_ticks (may be NULL), without enqueueing.  A block's result is bitwise the result
 * psdp_project_batched_f32 gives it alone.  cfg must auto-scale.  One host
 * thread drives a session (through its handle). */
typedef struct psdp_session_st* psdp_session_t;
psdp_status_t psdp_session_create(psdp_handle_t h, const psdp_config_t* cfg, size_t arena_bytes,
                                  psdp_session_t* session);
psdp_status_t psdp_session_destroy(psdp_session_t session);
/* Arena bytes a block of size n takes in this session. */
psdp_status_t psdp_session_block_bytes(psdp_session_t session, int64_t n, size_t* bytes);
psdp_status_t psdp_session_add(psdp_session_t session, const float* X, float* Y, int64_t n, int64_t* ticket);
psdp_status_t psdp_session_step(psdp_session_t session, int* in_flight_ticks);
psdp_status_t psdp_session_poll(psdp_session_t session, int64_t* tickets, int max, int* count, int* active,
                                int* in_flight_ticks);

/* on != 0 (the default): a product on 256 x 256 tiles whose last round fills
 * at most half the CUs splits that round's tiles' K range into chunks that run
 * side by side and are added in chunk order before the epilogue (no
 * counterpart in the reference: a tiling of this library; off, the round is
 * re-cut into 64 x 64 tiles).  Split tiles' entries differ from the whole-K
 * order in the last bits; with 0 every tile size gives bitwise the same
 * result. */
psdp_status_t psdp_set_split_k(psdp_handle_t h, int on);
/* Tile classes of the batched call: blocks with n >= big_block_n use
 * 128 x 128 tiles, the rest 64 x 64 (0: the built-in default); concurrent != 0
 * runs the two classes of each product on two streams. */
psdp_status_t psdp_set_batched_tiling(psdp_handle_t h, int big_block_n, int concurrent);
/* on != 0: the device-side sequence of a batched call (everything after the
 * descriptor upload) is captured once into a HIP graph and replayed by later
 * calls with the same block sizes, configuration, tiling and workspace. */
psdp_status_t psdp_set_batched_graph(psdp_handle_t h, int on);
/* Autotuner.  psdp_autotune times every tile size the product kernels offer
 * for a projection of size n with cfg (each operand format of its schedule,
 * on a synthetic symmetric matrix) and keeps the fastest in the handle; later
 * projections of that size use it.  psdp_tuned_tile reports the choice (tile
 * edge in elements, 0 if none).  psdp_autotune_batched runs the batched
 * projection of these blocks under each tile-class setting and keeps the
 * fastest (as psdp_set_batched_tiling would); it writes the blocks' outputs,
 * which every setting computes bitwise alike.  All block until done. */
psdp_status_t psdp_autotune(psdp_handle_t h, int64_t n, const psdp_config_t* cfg);
psdp_status_t psdp_tuned_tile(psdp_handle_t h, int64_t n, int format, int average, int* tile);
/* Sets the tile edge (64, 128 or 256; 0 clears) of products of size n in
 * `format` as the autotuner would; edges above the size's largest tile (256
 * from n = 2048 for 16-bit formats, 128 above 1024, else 64) are ignored. */
psdp_status_t psdp_set_tuned_tile(psdp_handle_t h, int64_t n, int format, int average, int tile);
psdp_status_t psdp_autotune_batched(psdp_handle_t h, const psdp_block_t* blocks, int count,
                                    const psdp_config_t* cfg);
/* Workspace the batched call needs (host only: also builds and checks every
 * launch descriptor of the call without touching a device). */
psdp_status_t psdp_workspace_size_batched(const psdp_block_t* blocks, int count, const psdp_config_t* cfg,
                                          size_t* bytes);

/* ---- EVD projections (rocSOLVER), drop-ins of eig_FP64_psd / eig_FP32_psd ----
 * eigenvalues: NULL or a caller-provided device array of n doubles that
 * receives the eigenvalues of the input (ascending) before clamping. */
psdp_status_t psdp_eig_fp64_psd(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues);
psdp_status_t psdp_eig_fp32_psd(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues);
/* The reference's ownership: with eigenvalues_out != NULL the library
 * allocates n doubles of device memory (hipMalloc), writes the eigenvalues
 * there and stores the pointer in *eigenvalues_out; the caller frees it with
 * hipFree or psdp_free (src/eig_FP64_psd.cu:46-49,84; the MEX caller frees it,
 * MATLAB/psd_projection_MATLAB.cu:158-162,175-179).  With eigenvalues_out ==
 * NULL nothing is allocated.  On failure *eigenvalues_out is NULL. */
psdp_status_t psdp_eig_fp64_psd_alloc(psdp_handle_t h, double* mat, int64_t n, double** eigenvalues_out);
psdp_status_t psdp_eig_fp32_psd_alloc(psdp_handle_t h, double* mat, int64_t n, double** eigenvalues_out);
/* Frees device memory the library allocated for the caller (NULL: no-op). */
psdp_status_t psdp_free(void* device_ptr);

/* ---- method-string front end (the MEX gateway's dispatch,
 * MATLAB/psd_projection_MATLAB.cu:112-166, method set README.md:56) ----
 * In-place float64 projection of mat by name: "composite_FP32",
 * "composite_FP16", "composite_FP32_emulated" (the *_auto_scale calls),
 * "eig_FP64", "eig_FP32" (EVD; eigenvalues: NULL or n device doubles that
 * receive the input's eigenvalues, as the MEX call returns them), or
 * "mixed" (the shipped schedule, auto-scaled).  Unknown names return
 * PSDP_ERR_INVALID_ARG.  Waits for the result, as the MEX call does. */
psdp_status_t psdp_project_method(psdp_handle_t h, const char* method, double* mat, int64_t n,
                                  double* eigenvalues);

/* ---- extremal eigenpairs and deflated projections (float64, device arrays) ---- */

/* Residual of Ritz pair j in psdp_compute_eigenpairs: |beta_last Z(row, j)|
 * for the eigenvector matrix Z of the Lanczos tridiagonal T. */
typedef enum psdp_ritz_residual {
  PSDP_RITZ_RESIDUAL_FIRST_ROW = 0, /* row 0, as src/lanczos.cu:295-301 reads it (default)       */
  PSDP_RITZ_RESIDUAL_LAST_ROW = 1   /* row nb-1: the Lanczos residual ||A x - theta x||          */
} psdp_ritz_residual_t;
/* Handle options of psdp_compute_eigenpairs: the residual row, and whether
 * the kept Ritz vectors Q z are divided by their norms (default 0: the
 * reference's unnormalised vectors).  The deflated FP32 projection uses the
 * residual row set here and always normalises. */
psdp_status_t psdp_set_eigenpair_options(psdp_handle_t h, int residual_row, int normalize_vectors);

/* compute_eigenpairs: Lanczos with the whole basis kept (max_iter = 0 -> n
 * iterations), Ritz pairs in order of decreasing |theta| kept while their
 * residual is < ortho_tol and their overlap with every kept vector is
 * <= ortho_tol, at most k.  eigenvalues (k) and eigenvectors (n x k, the
 * unnormalised Ritz vectors Q z) are device arrays; *r (host) receives the
 * number kept.  *upper_bound (host) = |theta_max| + residual + 1e-3, the
 * reference's return value.  upper_bound_only != 0 computes only the bound
 * (r, eigenvalues, eigenvectors may then be NULL).  tol is accepted and
 * unused, as in the reference.  Blocks until done. */
psdp_status_t psdp_compute_eigenpairs(psdp_handle_t h, const double* A, int64_t n, int64_t k, int64_t* r,
                                      double* eigenvalues, double* eigenvectors, int upper_bound_only,
                                      int64_t max_iter, double tol, double ortho_tol, double* upper_bound);

/* lobpcg: the m largest eigenpairs of A (3 m <= n).  V: n x m device array
 * (output; with warmstart != 0 also the orthonormal start block), D: m
 * device values in descending order (the reference documents D as m x m but
 * writes m values).  Stops when ||A X - X diag(D)||_F < tol or after maxiter
 * iterations (reference defaults: 100, 1e-8).  Blocks until done. */
psdp_status_t psdp_lobpcg(psdp_handle_t h, const double* A, double* V, double* D, int64_t n, int64_t m,
                          int warmstart, int maxiter, double tol);
/* Iterations and final residual norm of the handle's last lobpcg run. */
psdp_status_t psdp_lobpcg_last(psdp_handle_t h, int* iterations, double* residual_norm);

/* Deflated projections, in place (reference defaults: k = 40, tol = 1e-10,
 * ortho_tol = 1e-3, maxiter = 100).  Blocks until done.
 * fp32: the k largest-|lambda| Lanczos pairs (unit vectors) are removed, the rest is scaled
 * by s = 1.1 up + 1e-5 (up: Lanczos bound of the remainder) and projected by
 * the FP32 chain, the positive pairs are added back.
 * fp16 / eig_fp64: the k largest and the k smallest LOBPCG pairs (3 k <= n)
 * are removed, the rest is projected by the FP16 chain (after the same
 * scaling, from approximate_two_norm) or by the float64 EVD, the positive
 * pairs are added back.  ws/ws_bytes: optional workspace of the FP16 chain. */
psdp_status_t psdp_composite_fp32_auto_scale_deflate(psdp_handle_t h, double* mat, int64_t n, int64_t k,
                                                     double tol, double ortho_tol);
psdp_status_t psdp_composite_fp16_auto_scale_deflate(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                                     size_t ws_bytes, int64_t k, int maxiter, double tol);
psdp_status_t psdp_eig_fp64_deflate(psdp_handle_t h, double* mat, int64_t n, int64_t k, int maxiter, double tol);

#ifdef __cplusplus
}
#endif

#endif /* PSD_PROJECTION_PSDP_H */
