// Host driver and C ABI of libpsdp.
//
// A projection is a fixed sequence of launches on the handle's stream, with
// every scalar (the Lanczos bound and the scale) kept in device memory:
//   sym_pad (X -> symmetric float32, zero padded)
//   Lanczos bound on that matrix (auto-scaled entry points only)
//   scale_split (A0 = X/s, operand pieces of step 0 and of the final product)
//   3 symmetric products per filter step (fused polynomial epilogue)
//   final product W = X_s S with the output epilogue (x s, f64/f32, n x n)
// Workspace comes from the caller (ws, ws_bytes) or from the handle's arena,
// which grows on demand and is reused by later calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psd_projection/psdp.h"
#include "psdp_eigen.h"
#include "psdp_host.h"
#include "psdp_internal.h"

namespace psdp {
hipError_t launch_sym_product(int fmt, int sub, int avg, const ProdDesc& d, const TileRef* tiles, int ntiles,
                              hipStream_t s);
bool sym_big_tile_capable(int fmt, int avg);
hipError_t launch_split_reduce(int fmt, const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s);
hipError_t launch_sym_product_grouped(int fmt, int sub, int avg, const ProdDesc* descs, const TileRef* tiles,
                                      int ntiles, int grid, hipStream_t s);
hipError_t launch_sym_pad(const void* X, int x_f64, int64_t xld, int n, float* M, int64_t ld, int npad,
                          hipStream_t st);
hipError_t launch_scale_split(float* a32, int64_t count, int64_t ld, const double* inv, int fmt_a,
                              uint16_t* const ap[3],
                              int fmt_x, float* x32, uint16_t* const xp[3], hipStream_t st);
hipError_t launch_lanczos(const void* A, int a_f64, int64_t ld, int n, int m, int max_iter, double tol,
                          uint64_t seed, double margin, int tight, LanczosState* st, const LanczosVecs& v,
                          hipStream_t s);
hipError_t launch_set_scale(LanczosState* st, double scale, hipStream_t s);
hipError_t launch_fill_sym(float* M, int64_t ld, int n, int npad, uint64_t seed, hipStream_t st);
hipError_t launch_copy_words(void* dst, const void* src, int64_t words, hipStream_t st);
hipError_t launch_lanczos_batched(const LanczosJob* jobs, int count, int max_m, int max_iter, double tol,
                                  uint64_t seed, double margin, int tight, const TileRef* tiles, int ntiles,
                                  const int2* chunks, int nchunks, hipStream_t s);
hipError_t launch_sym_pad_batched(const PadJob* jobs, const TileRef* tiles, int ntiles, hipStream_t st);
hipError_t launch_scale_split_batched(const SplitJob* jobs, const int2* chunks, int nchunks, int fmt_a, int fmt_x,
                                      hipStream_t st);
size_t eig_workspace_bytes(int n, int f64);
psdp_status_t noalloc_scratch(int device, void** ptr, size_t* bytes);
hipError_t launch_fill_tile_list(TileRef* out, int T, hipStream_t s);
int eig_project(void* rb_handle, hipStream_t s, double* X, int n, int f64, void* ws, int* info_dev, double* eig);
void* eig_create_blas();
void eig_destroy_blas(void* rb);
}  // namespace psdp

using namespace psdp;

// ---------------------------------------------------------------- errors
static thread_local std::string g_last_error;

psdp_status_t psdp::fail(psdp_status_t st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

// ---------------------------------------------------------------- tables
// src/composite_FP32.cu:73-84 (also src/composite_FP32_emulated.cu:78-89)
static const double kTableFP32[10][3] = {
    {8.3119043343, -23.0739115930, 16.4664144722}, {4.1439360087, -2.9176674704, 0.5246212487},
    {4.0257813209, -2.9025002398, 0.5334261214},   {3.5118574347, -2.5740236523, 0.5050097282},
    {2.4398158400, -1.7586675341, 0.4191290613},   {1.9779835097, -1.3337358510, 0.3772169049},
    {1.9559726949, -1.3091355170, 0.3746734515},   {1.9282822454, -1.2823649693, 0.3704626545},
    {1.9220135179, -1.2812524618, 0.3707011753},   {1.8942192942, -1.2613293407, 0.3676616051}};
// src/composite_FP16.cu:93-101
static const double kTableFP16[7][3] = {
    {8.2885332412, -22.5927099246, 15.8201383114}, {4.1666196466, -2.9679004036, 0.5307623217},
    {4.0611848147, -2.9698947955, 0.5492133813},   {3.6678301399, -2.7561018955, 0.5421513305},
    {2.7632556383, -2.0607754898, 0.4695405857},   {2.0527445797, -1.4345145882, 0.4070669182},
    {1.8804816691, -1.2583997294, 0.3779501813}};

struct Schedule {
  int nsteps = 0;
  float a[PSDP_MAX_STEPS], b[PSDP_MAX_STEPS], c[PSDP_MAX_STEPS];
  int fmt[PSDP_MAX_STEPS];
  int avg[PSDP_MAX_STEPS];  // A^3 and A^5 products averaged with their transposes
  int final_fmt = FMT_F32;
  int final_avg = 0;
};

// The averaged (Jordan) product exists for the 1-piece 16-bit formats.
static int avg_capable(int fmt) { return fmt == FMT_F16 || fmt == FMT_BF16; }

// Coefficients as the reference computes them, in float arithmetic:
// a /= sf; b /= sf*sf*sf; c /= sf*sf*sf*sf*sf on the scaled steps
// (src/composite_FP32.cu:86-99, src/composite_FP16.cu:103-113).
static void fill_table(int table, Schedule& s) {
  const double(*t)[3] = table == PSDP_TABLE_FP16 ? kTableFP16 : kTableFP32;
  const int steps = table == PSDP_TABLE_FP16 ? 7 : 10;
  const int nscaled = table == PSDP_TABLE_FP16 ? 7 : 8;
  const float sf = table == PSDP_TABLE_FP16 ? 1.01f : 1.001f;
  s.nsteps = steps;
  for (int i = 0; i < steps; ++i) {
    float a = (float)t[i][0], b = (float)t[i][1], c = (float)t[i][2];
    if (i < nscaled) {
      const float f = sf;
      a /= f;
      b /= f * f * f;
      c /= f * f * f * f * f;
    }
    s.a[i] = a;
    s.b[i] = b;
    s.c[i] = c;
  }
}

// The reference symmetrizes A once per step by averaging with its transpose
// (src/utils.cu:7-32, called at src/composite_FP32.cu:118); with binary16
// operands the non-square products A2 A and A3 A2 are visibly non-symmetric,
// so the FP16 method averages them (C + C^T)/2 to keep the reference's error.
// Other formats keep the upper triangle (their commutator error is below the
// float32 accumulation error).
static psdp_status_t build_schedule(const psdp_config_t* cfg, Schedule& s) {
  for (int i = 0; i < PSDP_MAX_STEPS; ++i) s.avg[i] = 0;
  s.final_avg = 0;
  switch (cfg->method) {
    case PSDP_METHOD_COMPOSITE_FP32:
      fill_table(PSDP_TABLE_FP32, s);
      for (int i = 0; i < s.nsteps; ++i) s.fmt[i] = FMT_F32;
      s.final_fmt = FMT_F32;
      break;
    case PSDP_METHOD_COMPOSITE_FP16:
      fill_table(PSDP_TABLE_FP16, s);
      for (int i = 0; i < s.nsteps; ++i) s.fmt[i] = FMT_F16, s.avg[i] = 1;
      s.final_fmt = FMT_F16;
      s.final_avg = 1;
      break;
    case PSDP_METHOD_COMPOSITE_FP32_EMULATED:
      fill_table(PSDP_TABLE_FP32, s);
      for (int i = 0; i < s.nsteps; ++i) s.fmt[i] = FMT_BF16X9;
      s.final_fmt = FMT_BF16X9;
      break;
    case PSDP_METHOD_MIXED:
      // FP16-table chain: hi/lo split operands on steps 0-3, where rounding
      // errors are amplified by the later steps, plain binary16 on steps 4-6
      // and the final product, which only polish eigenvalues already near
      // +-1.  Measured (n = 640 .. 16384, three spectra): split products on
      // the final product change the error by < 1.2x; dropping them from any
      // of steps 1-3 costs 2-10x, from step 0 up to 3x when the scaled
      // spectrum reaches +-1.
      fill_table(PSDP_TABLE_FP16, s);
      for (int i = 0; i < s.nsteps; ++i) s.fmt[i] = i < 4 ? FMT_F16X3 : FMT_F16;
      s.final_fmt = FMT_F16;
      break;
    case PSDP_METHOD_MIXED_MX:
      // The mixed schedule with the split steps' cross terms on the scaled
      // FP8 MFMA (FMT_F16MX: two binary16-MFMA equivalents per k instead of
      // three); measured rel-F error 6.0-6.3e-4 against 5.1-5.3e-4 for mixed
      // (n = 1024-4096, three spectra, CPU emulation of both).
      fill_table(PSDP_TABLE_FP16, s);
      for (int i = 0; i < s.nsteps; ++i) s.fmt[i] = i < 4 ? FMT_F16MX : FMT_F16;
      s.final_fmt = FMT_F16;
      break;
    case PSDP_METHOD_CUSTOM:
      if (cfg->table != PSDP_TABLE_FP32 && cfg->table != PSDP_TABLE_FP16)
        return fail(PSDP_ERR_INVALID_ARG, "unknown coefficient table %d", cfg->table);
      fill_table(cfg->table, s);
      for (int i = 0; i < s.nsteps; ++i) {
        if (cfg->step_format[i] < 0 || cfg->step_format[i] >= FMT_COUNT)
          return fail(PSDP_ERR_INVALID_ARG, "step %d: unknown format %d", i, cfg->step_format[i]);
        s.fmt[i] = cfg->step_format[i];
        s.avg[i] = cfg->average_products && avg_capable(s.fmt[i]);
      }
      if (cfg->final_format < 0 || cfg->final_format >= FMT_COUNT)
        return fail(PSDP_ERR_INVALID_ARG, "unknown final format %d", cfg->final_format);
      s.final_fmt = cfg->final_format;
      s.final_avg = cfg->average_products && avg_capable(s.final_fmt);
      // FMT_F16MX pieces are written by the 16-bit direct-to-LDS kernels only
      for (int i = 0; i < s.nsteps; ++i) {
        const int fn = i + 1 < s.nsteps ? s.fmt[i + 1] : s.final_fmt;
        if (fn == FMT_F16MX && (s.fmt[i] == FMT_F32 || s.avg[i]))
          return fail(PSDP_ERR_UNSUPPORTED, "step %d: a float32 or averaged step cannot be followed by F16MX", i);
      }
      if (s.final_fmt == FMT_F16MX && s.final_avg) return fail(PSDP_ERR_UNSUPPORTED, "F16MX products are not averaged");
      break;
    default:
      return fail(PSDP_ERR_INVALID_ARG, "unknown method %d", cfg->method);
  }
  return PSDP_OK;
}

static int pieces_of(int fmt) {
  return fmt == FMT_F16X3 || fmt == FMT_BF16X3 ? 2 : (fmt == FMT_BF16X9 || fmt == FMT_F16MX ? 3 : (fmt == FMT_F32 ? 0 : 1));
}

// Bytes of piece q of an operand in format fmt, rows x ld 2-byte elements:
// FMT_F16MX's piece 2 holds one uint16 scale pair per row and 32-column
// block ((ld / 32) x ld of them).
static size_t piece_bytes(int fmt, int q, int64_t rows, int64_t ld) {
  if (q >= pieces_of(fmt)) return 0;
  if (fmt == FMT_F16MX && q == 2) return ((size_t)(ld / MX_BLOCK) * ld * 2 + 255) & ~(size_t)255;
  return ((size_t)rows * ld * 2 + 255) & ~(size_t)255;
}

// ---------------------------------------------------------------- handle
struct TileList {
  TileRef* dev = nullptr;
  int count = 0;
};

struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

// Launch lists of one product on 256 x 256 tiles (T x T of them, upper
// triangle): whole-K tiles, a re-cut 64 x 64 tail, and the K chunks of split
// tiles with the list their reduction runs over.
struct BigTiles {
  TileList big, small, split, reduce;
  int ksplit = 0, kchunk64 = 0;
};

struct psdp_handle_st {
  bool profiling = false;
  std::vector<ProfRec> prof;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  LanczosState* st = nullptr;  // device
  uint64_t seed = 0x5eed5eedULL;
  void* blas = nullptr;  // rocBLAS handle for the EVD drop-ins (created on first use)
  void* keep = nullptr;  // eigenpairs that outlive the projection of a deflated call
  size_t keep_bytes = 0;
  void* pinned = nullptr;  // small pinned host buffer for polled device scalars
  int ritz_row = PSDP_RITZ_RESIDUAL_FIRST_ROW;
  int ritz_normalize = 0;  // compute_eigenpairs returns the reference's unnormalised vectors by default
  int lobpcg_iters = 0;
  double lobpcg_rnorm = 0.0;
  // batched tile classes: blocks with n >= batch_big_n on 128 x 128 tiles;
  // with batch_concurrent the two classes of a stage run on two streams
  int batch_big_n = 0;  // 0: BIG_BLOCK_N
  std::map<std::tuple<int, int, int>, int> tuned;  // (n, format, averaged) -> tile sub (psdp_autotune)
  int batch_concurrent = 0;
  // states of the last batched call (in the arena) for psdp_last_bounds_batched
  const LanczosState* batch_states = nullptr;
  int batch_count = 0;
  hipEvent_t order_ev = nullptr;  // orders a new stream after work on the previous one
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // batched calls replayed from a captured HIP graph (psdp_set_batched_graph):
  // the graph of the last call and the key it was captured for
  int batch_graph = 0;
  hipStream_t cap = nullptr;
  hipGraphExec_t bgraph = nullptr;
  std::vector<int64_t> bgraph_key;
  std::map<std::pair<int, int>, TileList> tiles;  // (tiles per dim, tile size)
  std::map<std::pair<int, int>, struct BigTiles> big_tiles;  // (T, split products on) -> launch lists
  int split_k = 1;               // split products for under-filled rounds of 256 x 256 tiles (psdp_set_split_k)
  float* part = nullptr;         // partial tiles of split products
  size_t part_bytes = 0;
  int cus = 256;                                  // compute units of the device
  void* scratch = nullptr;                        // this device's scratch array (psdp_project_f32_noalloc)
  size_t scratch_bytes = 0;
  // pinned host staging ring for batched descriptors: slot k is reused only
  // after the copy recorded in ev[k] has run
  static constexpr int NSTAGE = 4;
  char* stage[NSTAGE] = {};
  size_t stage_bytes[NSTAGE] = {};
  hipEvent_t stage_ev[NSTAGE] = {};
  int stage_next = 0;
  std::mutex mu;
};

static psdp_status_t staging_slot(psdp_handle_t h, size_t bytes, char** out) {
  const int k = h->stage_next;
  if (h->stage_ev[k]) HIPCHK(hipEventSynchronize(h->stage_ev[k]));
  if (h->stage_bytes[k] < bytes) {
    if (h->stage[k]) HIPCHK(hipHostFree(h->stage[k]));
    h->stage[k] = nullptr;
    h->stage_bytes[k] = 0;
    if (hipHostMalloc((void**)&h->stage[k], bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PSDP_ERR_NOMEM, "pinned staging allocation of %zu bytes failed", bytes);
    }
    h->stage_bytes[k] = bytes;
  }
  if (!h->stage_ev[k]) HIPCHK(hipEventCreateWithFlags(&h->stage_ev[k], hipEventDisableTiming));
  *out = h->stage[k];
  return PSDP_OK;
}

// marks the current slot's copy (already enqueued on the handle's stream)
static psdp_status_t staging_release(psdp_handle_t h, hipStream_t on = nullptr) {
  const int k = h->stage_next;
  HIPCHK(hipEventRecord(h->stage_ev[k], on ? on : h->stream));
  h->stage_next = (k + 1) % psdp_handle_st::NSTAGE;
  return PSDP_OK;
}

// Upper-triangle tiles of a T x T grid, grouped into G x G super-tiles so that
// tiles running together share operand panels, then interleaved so that the
// hardware's round-robin dispatch gives each of the 8 XCDs one contiguous
// chunk of that order (its L2 then serves neighbouring panels).  G = 8 for
// 256 x 256 tiles: an XCD's 32 CUs then run 4 x 8 tiles of one super-tile (12
// distinct panels for 32 tiles instead of 16 with two 4 x 4 super-tiles);
// measured at n = 16384 (alternating runs on one device) a binary16x3 product
// takes 12.04-12.07 ms against 12.15-12.16 ms with G = 4.  Smaller tiles keep
// G = 4.
static std::vector<TileRef> super_tile_order(int T, int G = 4) {
  std::vector<TileRef> ord;
  ord.reserve((size_t)T * (T + 1) / 2);
  for (int si = 0; si < T; si += G)
    for (int sj = si; sj < T; sj += G)
      for (int i = si; i < std::min(si + G, T); ++i)
        for (int j = std::max(sj, i); j < std::min(sj + G, T); ++j) ord.push_back(TileRef{0, i, j, 0});
  return ord;
}

static std::vector<TileRef> xcd_interleave(const std::vector<TileRef>& ord) {
  const int L = (int)ord.size();
  const int XCD = 8;
  const int C = (L + XCD - 1) / XCD;
  std::vector<TileRef> out((size_t)C * XCD, TileRef{0, -1, -1, 0});
  for (int w = 0; w < C * XCD; ++w) {
    const int src = (w % XCD) * C + w / XCD;
    if (src < L) out[w] = ord[src];
  }
  return out;
}

static psdp_status_t upload_tiles(const std::vector<TileRef>& v, TileList& n) {
  n.count = (int)v.size();
  if (hipMalloc(&n.dev, std::max<size_t>(v.size(), 1) * sizeof(TileRef)) != hipSuccess)
    return fail(PSDP_ERR_NOMEM, "tile list allocation failed");
  if (!v.empty()) HIPCHK(hipMemcpy(n.dev, v.data(), v.size() * sizeof(TileRef), hipMemcpyHostToDevice));
  return PSDP_OK;
}

static psdp_status_t get_tiles(psdp_handle_t h, int T, int bt, TileList& tl) {
  auto key = std::make_pair(T, bt);
  auto it = h->tiles.find(key);
  if (it != h->tiles.end()) {
    tl = it->second;
    return PSDP_OK;
  }
  TileList n;
  STCHK(upload_tiles(xcd_interleave(super_tile_order(T)), n));
  h->tiles[key] = n;
  tl = n;
  return PSDP_OK;
}

// How a product on T x T tiles of 256 x 256 fills the device (cus CUs, one
// 256 x 256 workgroup each).  Whole rounds run whole-K tiles.  A last round
// that fills at most half the CUs is either re-cut into 64 x 64 tiles (two
// workgroups per CU, split products off; measured at n = 16384 a binary16x3
// product takes 12.14 ms with the 64 x 64 tail against 12.37 ms with a
// 128 x 128 tail) or, by default, split: each of its tiles runs as S K
// chunks (S = cus / tiles in the round, at most 16) whose partial sums a
// reduction adds in chunk order before the epilogue.  Measured per mixed_mx
// projection (ms, re-cut / split): n = 6144 12.15 / 11.32, 8192 23.79 /
// 23.84, 10240 42.0 / 40.4, 16384 156.1 / 154.8.  Splitting every tile of a
// launch with fewer tiles than CUs (S = 2-4) was slower (n = 4096: 5.09
// against 5.29 ms; per chunk the prologue, the partial tile's store and its
// reduction cost more than the idle CUs), so such launches run whole-K tiles.
// Split tiles sum their K in S pieces, so their entries differ from the
// whole-K order in the last bits (every other tiling agrees bitwise).  Chunks
// are whole multiples of 64 k, none empty.
struct SplitShape {
  int total = 0, full = 0, rem = 0;
  int nsplit = 0, S = 1, chunk64 = 0;  // split tiles (the last nsplit of the order), chunks each
  bool recut = false;                  // 64 x 64 tail instead (split products off)
};
static SplitShape split_shape(int T, int cus, bool split_on) {
  SplitShape z;
  const int ncu = std::max(cus, 1);
  z.total = T * (T + 1) / 2;
  z.full = z.total / ncu * ncu;
  z.rem = z.total - z.full;
  const int K64 = 4 * T;  // 64-element units of K (npad = 256 T)
  if (split_on && z.full > 0 && z.rem > 0 && 2 * z.rem <= ncu) {
    z.S = std::min(std::min(ncu / z.rem, 16), K64 / 2);
    if (z.S > 1) z.nsplit = z.rem;
  }
  if (z.nsplit) {
    z.chunk64 = (K64 + z.S - 1) / z.S;
    z.S = (K64 + z.chunk64 - 1) / z.chunk64;
  } else {
    z.S = 1;
  }
  z.recut = !z.nsplit && !(z.full == 0 || z.rem == 0 || 2 * z.rem > ncu);
  return z;
}

constexpr int TAIL_SUB = 2;  // 64 x 64
static psdp_status_t get_big_tiles(psdp_handle_t h, int T, const BigTiles** out) {
  const auto key = std::make_pair(T, h->split_k ? 1 : 0);
  auto it = h->big_tiles.find(key);
  if (it != h->big_tiles.end()) {
    *out = &it->second;
    return PSDP_OK;
  }
  const std::vector<TileRef> ord = super_tile_order(T, 8);
  const SplitShape z = split_shape(T, h->cus, h->split_k != 0);
  const int total = (int)ord.size();
  std::vector<TileRef> bigv, smallv, splitv, redv;
  if (z.recut) {
    bigv = xcd_interleave(std::vector<TileRef>(ord.begin(), ord.begin() + z.full));
    std::vector<TileRef> tail;
    constexpr int K = 8 / TAIL_SUB;  // tail tiles per big tile edge
    for (int t = z.full; t < total; ++t)
      for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) {
          const int r = K * ord[t].bi + a, c = K * ord[t].bj + b;
          if (r <= c) tail.push_back(TileRef{0, r, c, 0});  // the lower part of a diagonal tile is a mirror
        }
    smallv = xcd_interleave(tail);
  } else {
    const int nbig = total - z.nsplit;
    if (nbig) bigv = xcd_interleave(std::vector<TileRef>(ord.begin(), ord.begin() + nbig));
    std::vector<TileRef> chunks;
    for (int s = 0; s < (z.nsplit ? z.S : 0); ++s)
      for (int j = 0; j < z.nsplit; ++j) chunks.push_back(TileRef{j, ord[nbig + j].bi, ord[nbig + j].bj, s});
    if (!chunks.empty()) splitv = xcd_interleave(chunks);
    for (int j = 0; j < z.nsplit; ++j) redv.push_back(TileRef{j, ord[nbig + j].bi, ord[nbig + j].bj, 0});
  }
  BigTiles bt;
  auto up = [&](const std::vector<TileRef>& v, TileList& l) -> psdp_status_t {
    if (v.empty()) return PSDP_OK;
    return upload_tiles(v, l);
  };
  STCHK(up(bigv, bt.big));
  STCHK(up(smallv, bt.small));
  STCHK(up(splitv, bt.split));
  STCHK(up(redv, bt.reduce));
  bt.ksplit = z.nsplit ? z.S : 0;
  bt.kchunk64 = z.chunk64;
  auto& slot = h->big_tiles[key];
  slot = bt;
  *out = &slot;
  return PSDP_OK;
}

// Device memory for the partial tiles of split products (grown on demand;
// the products of one stream use it one after another).
static psdp_status_t part_reserve(psdp_handle_t h, size_t bytes, float** out) {
  if (bytes > h->part_bytes) {
    if (h->part) {
      HIPCHK(hipStreamSynchronize(h->stream));
      HIPCHK(hipFree(h->part));
      h->part = nullptr;
      h->part_bytes = 0;
    }
    HIPCHK(hipMalloc((void**)&h->part, bytes));
    h->part_bytes = bytes;
  }
  *out = h->part;
  return PSDP_OK;
}

// One product on 256 x 256 tiles: whole-K tiles, then the re-cut tail or the
// split tiles' chunks and their reduction (which runs the epilogue).
static psdp_status_t launch_big_product(psdp_handle_t h, int T, int fmt, int avg, const ProdDesc& d, hipStream_t s) {
  const BigTiles* bt = nullptr;
  STCHK(get_big_tiles(h, T, &bt));
  if (bt->big.count) HIPCHK(launch_sym_product(fmt, 8, avg, d, bt->big.dev, bt->big.count, s));
  if (bt->small.count) HIPCHK(launch_sym_product(fmt, TAIL_SUB, avg, d, bt->small.dev, bt->small.count, s));
  if (bt->split.count) {
    float* part = nullptr;
    STCHK(part_reserve(h, (size_t)bt->reduce.count * bt->ksplit * 65536 * sizeof(float), &part));
    ProdDesc dp = d;
    dp.mode = EPI_PARTIAL;
    dp.part = part;
    dp.ksplit = bt->ksplit;
    dp.kchunk64 = bt->kchunk64;
    HIPCHK(launch_sym_product(fmt, 8, avg, dp, bt->split.dev, bt->split.count, s));
    ProdDesc dr = d;
    dr.part = part;
    dr.ksplit = bt->ksplit;
    dr.kchunk64 = bt->kchunk64;
    HIPCHK(launch_split_reduce(fmt, dr, bt->reduce.dev, bt->reduce.count, s));
  }
  return PSDP_OK;
}

static psdp_status_t arena_reserve(psdp_handle_t h, size_t bytes, void** out) {
  h->batch_states = nullptr;  // the caller overwrites the arena: the last batch's bounds go
  h->batch_count = 0;
  if (bytes > h->arena_bytes) {
    if (h->arena) {
      HIPCHK(hipStreamSynchronize(h->stream));
      HIPCHK(hipFree(h->arena));
      h->arena = nullptr;
      h->arena_bytes = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PSDP_ERR_NOMEM, "workspace allocation of %zu bytes failed", bytes);
    }
    h->arena = p;
    h->arena_bytes = bytes;
  }
  *out = h->arena;
  return PSDP_OK;
}

// Brackets one launch with events when profiling is on.
struct ProfScope {
  psdp_handle_t h;
  ProfRec rec;
  bool on;
  ProfScope(psdp_handle_t hh, int kind) : h(hh), on(hh->profiling) {
    if (!on) return;
    rec.kind = kind;
    if (hipEventCreate(&rec.a) != hipSuccess || hipEventCreate(&rec.b) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(rec.a, h->stream);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.b, h->stream);
    h->prof.push_back(rec);
  }
};

static constexpr size_t kNone = ~(size_t)0;

// Host-side check of a product descriptor before it is launched: every
// operand piece the format reads and every buffer the epilogue writes must be
// present, so that a planning mistake becomes an error instead of a fault.
static bool desc_ok(const ProdDesc& d, int fmt, const char** why) {
  const int np = fmt == FMT_F32 ? 1 : pieces_of(fmt);
  for (int i = 0; i < np; ++i)
    if (!d.x[i] || !d.y[i]) return *why = "missing operand piece", false;
  if (d.npad <= 0 || d.npad % 64 || d.ld < d.npad || d.n > d.npad) return *why = "bad sizes", false;
  if ((fmt == FMT_F16MX || d.out_fmt == FMT_F16MX) && d.ld % MX_BLOCK) return *why = "ld not a multiple of 32", false;
  switch (d.mode) {
    case EPI_STORE:
      if (!d.c32 && d.out_fmt == FMT_F32) return *why = "no float32 destination", false;
      break;
    case EPI_POLY:
    case EPI_POLY_HALF_ID:
      // binary16x3 steps read A^3 back from the pieces of X (p3 null)
      if (!d.c32 || (!d.p3 && fmt != FMT_F16X3)) return *why = "missing A or A^3", false;
      break;
    case EPI_OUTPUT:
      if (!d.out || d.out_ld < d.n) return *why = "missing output", false;
      return true;
    default:
      return *why = "bad mode", false;
  }
  if (d.out_fmt != FMT_F32) {
    const int no = pieces_of(d.out_fmt);
    for (int i = 0; i < no; ++i)
      if (!d.outp[i]) return *why = "missing output piece", false;
  }
  return true;
}

#define DESCCHK(d, fmt)                                                                   \
  do {                                                                                    \
    const char* why_ = "";                                                                \
    if (!desc_ok((d), (fmt), &why_)) return fail(PSDP_ERR_INVALID_ARG, "internal: %s", why_); \
  } while (0)

// ---------------------------------------------------------------- plan
// Row stride of the workspace matrices: npad + LD_PAD elements.  A
// power-of-two stride (npad = 16384 is 32 KB per binary16 row) maps every row
// of a tile panel to the same cache sets and channels; 64 extra elements
// (128 B for binary16, 256 B for float32) spread them.
static constexpr int LD_PAD = 64;

struct Plan {
  int n = 0, npad = 0, ld = 0, sub = 4, bt = 128;
  int64_t N = 0;
  Schedule sch;
  int lanczos_iters = 20;
  size_t off_a32 = 0, off_p3 = 0, off_p2 = kNone, off_x32 = kNone;
  size_t off_ap[3] = {kNone, kNone, kNone}, off_p2p[3] = {kNone, kNone, kNone};
  size_t off_p3p[3] = {kNone, kNone, kNone}, off_xp[3] = {kNone, kNone, kNone};
  size_t off_vec = 0;
  size_t total = 0;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Piece sizes of the iterate A (every step format and the final one) and of
// A^2 / A^3 (step formats), per piece index.
static void piece_sizes(const Schedule& sch, int64_t rows, int64_t ld, size_t ab[3], size_t sb[3]) {
  for (int q = 0; q < 3; ++q) {
    ab[q] = piece_bytes(sch.final_fmt, q, rows, ld);
    sb[q] = 0;
    for (int i = 0; i < sch.nsteps; ++i) {
      ab[q] = std::max(ab[q], piece_bytes(sch.fmt[i], q, rows, ld));
      sb[q] = std::max(sb[q], piece_bytes(sch.fmt[i], q, rows, ld));
    }
  }
}

static int choose_sub(int64_t n) { return n > 1024 ? 4 : 2; }

// Largest tile a product of size n may use (the plan pads n to a multiple of
// it, so every smaller tile divides the padded size too): 256 x 256 from
// n = 2048 where the kernel exists, else 128 (n > 1024) or 64.
static int max_sub(int64_t n, int fmt, int avg) {
  if (n >= 2048 && sym_big_tile_capable(fmt, avg)) return 8;
  return choose_sub(n);
}

// Built-in tile choice (psdp_autotune can override it per size) for the
// 16-bit direct-to-LDS kernels: the tile size with the smallest modelled
// launch cost, (rounds of tiles over the device's workgroup slots) x (time
// of one round relative to a round of 64 x 64 tiles: 1.65 for 128 x 128,
// 4.3 for 256 x 256, each over the same K).  64 x 64 tiles run two
// workgroups per CU, the larger ones one; a partial last round of 256 x 256
// tiles that fills at most half the CUs is re-cut into 64 x 64 tiles
// (get_split_tiles).  The weights come from timing every tile size on
// mixed and mixed_mx projections (ms, 64 / 128 / 256; all results bitwise
// equal): n = 2560: 2.16 / 2.00 / 3.58, 3072: 3.02 / 3.23 / 4.00, 3584:
// 4.30 / 3.75 / 4.51, 4096: 5.79 / 5.57 / 5.08, 5120: 10.1 / 9.0 / 6.9,
// 6144: 16.8 / 13.8 / 12.0, 7168: 25.8 / 21.3 / 17.0, 8192: 41.4 / 31.5 /
// 23.6; the model picks the fastest at every one of them.  Other formats
// keep 128 x 128 above n = 1024.
static int launch_sub(int64_t n, int fmt, int avg, int cus) {
  if (!sym_big_tile_capable(fmt, avg)) return choose_sub(n);
  const int top = max_sub(n, fmt, avg);
  const int64_t btop = 32 * top, npad = (n + btop - 1) / btop * btop;
  const int64_t slots = std::max(cus, 1);
  auto tiles = [&](int64_t bt) {
    const int64_t T = npad / bt;
    return T * (T + 1) / 2;
  };
  auto rounds = [](int64_t t, int64_t per) { return (double)((t + per - 1) / per); };
  int sub = 2;
  double best = rounds(tiles(64), 2 * slots);
  if (top >= 4) {
    const double c = 1.65 * rounds(tiles(128), slots);
    if (c < best) best = c, sub = 4;
  }
  if (top >= 8) {
    const int64_t total = tiles(256), full = total / slots * slots, rem = total - full;
    const double c = (full == 0 || rem == 0 || 2 * rem > slots)
                         ? 4.3 * rounds(total, slots)
                         : 4.3 * (double)(full / slots) + rounds(16 * rem, 2 * slots);
    if (c < best) best = c, sub = 8;
  }
  return sub;
}

static psdp_status_t make_plan(const psdp_config_t* cfg, int64_t n, Plan& p) {
  if (n <= 0) return fail(PSDP_ERR_INVALID_ARG, "matrix size must be positive (got %lld)", (long long)n);
  if (n > (1 << 20)) return fail(PSDP_ERR_INVALID_ARG, "matrix size %lld too large", (long long)n);
  STCHK(build_schedule(cfg, p.sch));
  p.lanczos_iters = cfg->lanczos_iters > 0 ? cfg->lanczos_iters : 20;
  if (p.lanczos_iters > LANCZOS_MAX_ITER)
    return fail(PSDP_ERR_INVALID_ARG, "at most %d Lanczos iterations are supported", LANCZOS_MAX_ITER);
  p.n = (int)n;
  p.sub = choose_sub(n);
  for (int i = 0; i < p.sch.nsteps; ++i) {
    p.sub = std::max(p.sub, max_sub(n, p.sch.fmt[i], 0));
    p.sub = std::max(p.sub, max_sub(n, p.sch.fmt[i], p.sch.avg[i]));
  }
  p.sub = std::max(p.sub, max_sub(n, p.sch.final_fmt, p.sch.final_avg));
  p.bt = 32 * p.sub;  // the largest tile used: npad is a multiple of every tile
  p.npad = (int)((n + p.bt - 1) / p.bt * p.bt);
  p.ld = p.npad + LD_PAD;
  p.N = (int64_t)p.npad * p.ld;
  bool any_f32 = false;
  int pa = pieces_of(p.sch.final_fmt), pp = 0;
  for (int i = 0; i < p.sch.nsteps; ++i) {
    any_f32 |= p.sch.fmt[i] == FMT_F32;
    pa = std::max(pa, pieces_of(p.sch.fmt[i]));
    pp = std::max(pp, pieces_of(p.sch.fmt[i]));
  }
  const int px = pieces_of(p.sch.final_fmt);
  const size_t f32b = align256((size_t)p.N * 4);
  size_t off = 0;
  p.off_a32 = off, off += f32b;
  p.off_p3 = off, off += f32b;
  if (any_f32) p.off_p2 = off, off += f32b;
  if (p.sch.final_fmt == FMT_F32) p.off_x32 = off, off += f32b;
  size_t ab[3], sb[3];
  piece_sizes(p.sch, p.npad, p.ld, ab, sb);
  for (int i = 0; i < pa; ++i) p.off_ap[i] = off, off += ab[i];
  for (int i = 0; i < pp; ++i) p.off_p2p[i] = off, off += sb[i];
  for (int i = 0; i < pp; ++i) p.off_p3p[i] = off, off += sb[i];
  for (int i = 0; i < px; ++i) p.off_xp[i] = off, off += piece_bytes(p.sch.final_fmt, i, p.npad, p.ld);
  p.off_vec = off;
  off += align256((size_t)p.npad * 8) * 6 + align256((size_t)p.npad * 8 * p.lanczos_iters);
  p.total = off;
  return PSDP_OK;
}

struct Bufs {
  float *a32, *p3, *p2, *x32;
  uint16_t *ap[3], *p2p[3], *p3p[3], *xp[3];
  LanczosVecs v;
};

static Bufs carve(const Plan& p, char* base) {
  Bufs b;
  b.a32 = (float*)(base + p.off_a32);
  b.p3 = (float*)(base + p.off_p3);
  b.p2 = p.off_p2 != kNone ? (float*)(base + p.off_p2) : nullptr;
  b.x32 = p.off_x32 != kNone ? (float*)(base + p.off_x32) : nullptr;
  for (int i = 0; i < 3; ++i) {
    b.ap[i] = p.off_ap[i] != kNone ? (uint16_t*)(base + p.off_ap[i]) : nullptr;
    b.p2p[i] = p.off_p2p[i] != kNone ? (uint16_t*)(base + p.off_p2p[i]) : nullptr;
    b.p3p[i] = p.off_p3p[i] != kNone ? (uint16_t*)(base + p.off_p3p[i]) : nullptr;
    b.xp[i] = p.off_xp[i] != kNone ? (uint16_t*)(base + p.off_xp[i]) : nullptr;
  }
  const size_t vb = align256((size_t)p.npad * 8);
  char* v = base + p.off_vec;
  b.v.q = (double*)(v + 0 * vb);
  b.v.w = (double*)(v + 1 * vb);
  b.v.w1 = (double*)(v + 2 * vb);
  b.v.vold = (double*)(v + 3 * vb);
  b.v.y = (double*)(v + 4 * vb);
  b.v.ay = (double*)(v + 5 * vb);
  b.v.V = (double*)(v + 6 * vb);
  // the Lanczos GEMV's partial sums (ceil(npad/256) x npad doubles, at most
  // half of npad x ld floats) use A^3's buffer, which is written only later
  b.v.part = (double*)b.p3;
  return b;
}

static void set_operand(const void* (&dst)[3], int fmt, const float* m32, uint16_t* const pieces[3]) {
  if (fmt == FMT_F32) {
    dst[0] = m32;
    dst[1] = dst[2] = nullptr;
  } else {
    for (int i = 0; i < 3; ++i) dst[i] = pieces[i];
  }
}

static ProdDesc base_desc(const Plan& p) {
  ProdDesc d;
  memset(&d, 0, sizeof d);
  d.ld = p.ld;
  d.npad = p.npad;
  d.n = p.n;
  return d;
}

// Tile lists that live in caller-provided device memory (the allocation-free
// entry point builds them there), by tile sub-size.
struct FixedTiles {
  const TileRef* list[9] = {};
  int count[9] = {};
};

// Enqueues the whole projection.
// One symmetric product on tiles of 32 * sub: sub 8 runs whole rounds of
// 256 x 256 tiles and re-cuts a partial last round into 128 x 128 tiles.
static psdp_status_t launch_product(psdp_handle_t h, const Plan& p, int fmt, int avg, const ProdDesc& d, int sub,
                                    hipStream_t s, const FixedTiles* ft = nullptr) {
  if (ft) {
    if (sub < 0 || sub > 8 || !ft->list[sub]) return fail(PSDP_ERR_INVALID_ARG, "internal: no tile list for %d", sub);
    HIPCHK(launch_sym_product(fmt, sub, avg, d, ft->list[sub], ft->count[sub], s));
    return PSDP_OK;
  }
  if (sub == 8) return launch_big_product(h, p.npad / 256, fmt, avg, d, s);
  TileList tl;
  STCHK(get_tiles(h, p.npad / (32 * sub), 32 * sub, tl));
  HIPCHK(launch_sym_product(fmt, sub, avg, d, tl.dev, tl.count, s));
  return PSDP_OK;
}

// Tile size of a product: the autotuner's choice for (n, format, averaged)
// when psdp_autotune measured one, else the built-in rule.  Choices never
// exceed the built-in tile, so the plan's padding holds for all of them.
static int tuned_sub(psdp_handle_t h, const Plan& p, int fmt, int avg) {
  auto it = h->tuned.find(std::make_tuple(p.n, fmt, avg));
  if (it != h->tuned.end() && it->second <= max_sub(p.n, fmt, avg)) return it->second;
  return launch_sub(p.n, fmt, avg, h->cus);
}

// fixed_scale: s of the entry points without the bound (1 for the reference's
// composite_* functions, the deflated matrix's scale for the *_deflate ones).
// st: device Lanczos state (default: the handle's); ft: tile lists in caller
// memory (default: the handle's cached lists).
// Whether a step in operand format fmt keeps a float32 copy of A^3 for its
// polynomial epilogue.  A binary16x3 step does not: the epilogue reads A^3
// back from its two pieces (within 2^-22 |v| + 2^-33 of the float32 value v,
// the precision the products already take it at), which saves the A^3 product a
// float32 store of every entry.  Other formats keep the copy (one binary16 or
// bfloat16 piece is too coarse; the float32, FP8 and MX formats have no
// pieces to read it from).
static bool a3_f32(int fmt) { return fmt != FMT_F16X3; }

static psdp_status_t enqueue_projection(psdp_handle_t h, const Plan& p, const Bufs& b, const void* X, int x_f64,
                                        void* Y, int y_f64, const psdp_config_t* cfg, double fixed_scale = 1.0,
                                        LanczosState* st = nullptr, const FixedTiles* ft = nullptr) {
  hipStream_t s = h->stream;
  const Schedule& sch = p.sch;
  if (!st) st = h->st;
  // one product: tile size by format, then the launch
  auto product = [&](int fmt, int avg, const ProdDesc& d) -> psdp_status_t {
    DESCCHK(d, fmt);
    ProfScope ps(h, PSDP_PROF_PROD_F32 + fmt);
    return launch_product(h, p, fmt, avg, d, tuned_sub(h, p, fmt, avg), s, ft);
  };

  {
    ProfScope ps(h, PSDP_PROF_PREP);
    HIPCHK(launch_sym_pad(X, x_f64, p.n, p.n, b.a32, p.ld, p.npad, s));
  }
  if (cfg->auto_scale) {
    const double tol = cfg->lanczos_tol > 0 ? cfg->lanczos_tol : 1e-10;
    const double margin = cfg->scale_margin > 0 ? cfg->scale_margin : 1.0;
    const int iters = std::min(p.lanczos_iters, p.n);  // max_iter = min(max_iter, n), src/lanczos.cu:55
    ProfScope ps(h, PSDP_PROF_LANCZOS);
    HIPCHK(launch_lanczos(b.a32, 0, p.ld, p.n, p.npad, iters, tol, h->seed, margin,
                          cfg->bound_mode == PSDP_BOUND_TIGHT, st, b.v, s));
  } else {
    HIPCHK(launch_set_scale(st, fixed_scale, s));
  }
  {
    ProfScope ps(h, PSDP_PROF_PREP);
    HIPCHK(launch_scale_split(b.a32, p.N, p.ld, &st->inv_scale, sch.fmt[0], b.ap, sch.final_fmt, b.x32, b.xp, s));
  }

  for (int i = 0; i < sch.nsteps; ++i) {
    const int f = sch.fmt[i];
    const int fn = i + 1 < sch.nsteps ? sch.fmt[i + 1] : sch.final_fmt;
    const bool last = i + 1 == sch.nsteps;
    // A2 = A A
    ProdDesc d1 = base_desc(p);
    set_operand(d1.x, f, b.a32, b.ap);
    set_operand(d1.y, f, b.a32, b.ap);
    d1.mode = EPI_STORE;
    d1.c32 = f == FMT_F32 ? b.p2 : nullptr;
    d1.out_fmt = f;
    for (int q = 0; q < 3; ++q) d1.outp[q] = b.p2p[q];
    STCHK(product(f, 0, d1));
    // A3 = A2 A
    ProdDesc d2 = base_desc(p);
    set_operand(d2.x, f, b.p2, b.p2p);
    set_operand(d2.y, f, b.a32, b.ap);
    d2.mode = EPI_STORE;
    d2.c32 = a3_f32(f) ? b.p3 : nullptr;
    d2.out_fmt = f;
    for (int q = 0; q < 3; ++q) d2.outp[q] = b.p3p[q];
    STCHK(product(f, sch.avg[i], d2));
    // A = a A + b A3 + c A3 A2 (last step: S = (I + A)/2)
    ProdDesc d3 = base_desc(p);
    set_operand(d3.x, f, b.p3, b.p3p);
    set_operand(d3.y, f, b.p2, b.p2p);
    d3.mode = last ? EPI_POLY_HALF_ID : EPI_POLY;
    d3.a = sch.a[i];
    d3.b = sch.b[i];
    d3.c = sch.c[i];
    d3.c32 = b.a32;
    d3.p3 = a3_f32(f) ? b.p3 : nullptr;
    d3.out_fmt = fn;
    for (int q = 0; q < 3; ++q) d3.outp[q] = b.ap[q];
    STCHK(product(f, sch.avg[i], d3));
  }
  // W = X_s S, Y = s W
  ProdDesc d4 = base_desc(p);
  set_operand(d4.x, sch.final_fmt, b.x32, b.xp);
  set_operand(d4.y, sch.final_fmt, b.a32, b.ap);
  d4.mode = EPI_OUTPUT;
  d4.out = Y;
  d4.out_ld = p.n;
  d4.out_f64 = y_f64;
  d4.scale = &st->scale;
  d4.nonfinite = &st->out_nonfinite;
  STCHK(product(sch.final_fmt, sch.final_avg, d4));
  return PSDP_OK;
}

static void clear_profile(psdp_handle_t h) {
  for (auto& r : h->prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  h->prof.clear();
}

static psdp_status_t check_flags(psdp_handle_t h, const LanczosState* dev_state = nullptr) {
  LanczosState st;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&st, dev_state ? dev_state : h->st, sizeof st, hipMemcpyDeviceToHost));
  if (st.nonfinite) return fail(PSDP_ERR_NONFINITE, "non-finite value in the input or the spectral bound");
  if (st.out_nonfinite)
    return fail(PSDP_ERR_NONFINITE,
                "non-finite value in the result: the filter chain diverged (a spectrum outside [-1, 1] "
                "needs the auto-scaled entry point)");
  return PSDP_OK;
}

// Caller holds h->mu and has made h->device current.
static psdp_status_t project_locked(psdp_handle_t h, const void* X, int x_f64, void* Y, int y_f64, int64_t n,
                                    const psdp_config_t* cfg, void* ws, size_t ws_bytes, double fixed_scale = 1.0) {
  Plan p;
  STCHK(make_plan(cfg, n, p));
  void* base = ws;
  if (ws) {
    if (ws_bytes < p.total)
      return fail(PSDP_ERR_INVALID_ARG, "workspace of %zu bytes is smaller than the %zu bytes needed", ws_bytes,
                  p.total);
  } else {
    STCHK(arena_reserve(h, p.total, &base));
  }
  Bufs b = carve(p, (char*)base);
  STCHK(enqueue_projection(h, p, b, X, x_f64, Y, y_f64, cfg, fixed_scale));
  if (cfg->synchronize) STCHK(check_flags(h));
  return PSDP_OK;
}

static psdp_status_t project(psdp_handle_t h, const void* X, int x_f64, void* Y, int y_f64, int64_t n,
                             const psdp_config_t* cfg, void* ws, size_t ws_bytes) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (!X || !Y) return fail(PSDP_ERR_INVALID_ARG, "null matrix pointer");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  return project_locked(h, X, x_f64, Y, y_f64, n, cfg, ws, ws_bytes);
}

// ---------------------------------------------------------------- C ABI
extern "C" {

const char* psdp_last_error(void) { return g_last_error.c_str(); }

const char* psdp_status_string(psdp_status_t s) {
  switch (s) {
    case PSDP_OK: return "PSDP_OK";
    case PSDP_ERR_INVALID_ARG: return "PSDP_ERR_INVALID_ARG";
    case PSDP_ERR_NOMEM: return "PSDP_ERR_NOMEM";
    case PSDP_ERR_NONFINITE: return "PSDP_ERR_NONFINITE";
    case PSDP_ERR_UNSUPPORTED: return "PSDP_ERR_UNSUPPORTED";
    case PSDP_ERR_DEVICE: return "PSDP_ERR_DEVICE";
  }
  return "PSDP_ERR_UNKNOWN";
}

psdp_status_t psdp_default_config(psdp_config_t* cfg, int method) {
  if (!cfg) return fail(PSDP_ERR_INVALID_ARG, "null config");
  memset(cfg, 0, sizeof *cfg);
  cfg->method = method;
  cfg->table = PSDP_TABLE_FP32;
  for (int i = 0; i < PSDP_MAX_STEPS; ++i) cfg->step_format[i] = PSDP_FMT_F32;
  cfg->final_format = PSDP_FMT_F32;
  cfg->auto_scale = 1;
  cfg->lanczos_iters = 20;
  cfg->lanczos_tol = 1e-10;
  cfg->scale_margin = 1.0;
  cfg->synchronize = 1;
  return PSDP_OK;
}

psdp_status_t psdp_create(psdp_handle_t* out, int device) {
  if (!out) return fail(PSDP_ERR_INVALID_ARG, "null handle pointer");
  *out = nullptr;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(PSDP_ERR_INVALID_ARG, "device %d not present", device);
  HIPCHK(hipSetDevice(device));
  psdp_handle_t h = new psdp_handle_st();
  h->device = device;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(PSDP_ERR_DEVICE, "stream creation failed");
  }
  h->own_stream = true;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    h->cus = cus;
  if (hipMalloc(&h->st, sizeof(LanczosState)) != hipSuccess) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return fail(PSDP_ERR_NOMEM, "state allocation failed");
  }
  (void)hipMemset(h->st, 0, sizeof(LanczosState));
  // the device address of this device's scratch array, looked up once per
  // device with the device current (the allocation-free entry point uses it)
  const psdp_status_t sst = noalloc_scratch(device, &h->scratch, &h->scratch_bytes);
  if (sst != PSDP_OK) {
    (void)hipFree(h->st);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return sst;
  }
  *out = h;
  return PSDP_OK;
}

psdp_status_t psdp_profile_enable(psdp_handle_t h, int on) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipStreamSynchronize(h->stream));
  if (on) clear_profile(h);
  h->profiling = on != 0;
  return PSDP_OK;
}

psdp_status_t psdp_profile_summary(psdp_handle_t h, int kind, int* launches, double* total_ms) {
  if (!h || !launches || !total_ms) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipStreamSynchronize(h->stream));
  int c = 0;
  double t = 0.0;
  for (auto& r : h->prof) {
    if (r.kind != kind) continue;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    ++c;
    t += ms;
  }
  *launches = c;
  *total_ms = t;
  return PSDP_OK;
}

psdp_status_t psdp_destroy(psdp_handle_t h) {
  if (!h) return PSDP_OK;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  clear_profile(h);
  for (auto& kv : h->tiles) (void)hipFree(kv.second.dev);
  for (auto& kv : h->big_tiles)
    for (TileList* l : {&kv.second.big, &kv.second.small, &kv.second.split, &kv.second.reduce})
      if (l->dev) (void)hipFree(l->dev);
  if (h->part) (void)hipFree(h->part);
  for (int k = 0; k < psdp_handle_st::NSTAGE; ++k) {
    if (h->stage_ev[k]) (void)hipEventSynchronize(h->stage_ev[k]);
    if (h->stage_ev[k]) (void)hipEventDestroy(h->stage_ev[k]);
    if (h->stage[k]) (void)hipHostFree(h->stage[k]);
  }
  if (h->arena) (void)hipFree(h->arena);
  if (h->keep) (void)hipFree(h->keep);
  if (h->bgraph) (void)hipGraphExecDestroy(h->bgraph);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
  if (h->order_ev) (void)hipEventDestroy(h->order_ev);
  if (h->join_ev) (void)hipEventDestroy(h->join_ev);
  if (h->aux) (void)hipStreamDestroy(h->aux);
  if (h->pinned) (void)hipHostFree(h->pinned);
  if (h->st) (void)hipFree(h->st);
  eig_destroy_blas(h->blas);
  if (h->own_stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PSDP_OK;
}

// Work already enqueued on the previous stream may still read or write the
// handle's arena, Lanczos state and staging memory, which the next call
// reuses: the new stream waits for an event recorded on the old one (no host
// wait), so calls through one handle stay ordered across stream changes.
psdp_status_t psdp_set_stream(psdp_handle_t h, void* stream) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  hipStream_t ns = (hipStream_t)stream;
  if (ns == h->stream && !h->own_stream) return PSDP_OK;
  HIPCHK(hipSetDevice(h->device));
  if (h->own_stream) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipStreamDestroy(h->stream));
    h->own_stream = false;
  } else {
    if (!h->order_ev) HIPCHK(hipEventCreateWithFlags(&h->order_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(h->order_ev, h->stream));
    HIPCHK(hipStreamWaitEvent(ns, h->order_ev, 0));
  }
  h->stream = ns;
  return PSDP_OK;
}

psdp_status_t psdp_get_stream(psdp_handle_t h, void** stream) {
  if (!h || !stream) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  *stream = (void*)h->stream;
  return PSDP_OK;
}

psdp_status_t psdp_set_seed(psdp_handle_t h, uint64_t seed) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  h->seed = seed;
  return PSDP_OK;
}

psdp_status_t psdp_set_batched_graph(psdp_handle_t h, int on) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  h->batch_graph = on != 0;
  return PSDP_OK;
}

psdp_status_t psdp_set_split_k(psdp_handle_t h, int on) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  h->split_k = on != 0;
  return PSDP_OK;
}

psdp_status_t psdp_set_batched_tiling(psdp_handle_t h, int big_block_n, int concurrent) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (big_block_n < 0) return fail(PSDP_ERR_INVALID_ARG, "negative block size");
  std::lock_guard<std::mutex> lock(h->mu);
  h->batch_big_n = big_block_n;
  h->batch_concurrent = concurrent != 0;
  return PSDP_OK;
}

psdp_status_t psdp_synchronize(psdp_handle_t h) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return PSDP_OK;
}

// ---------------------------------------------------------------- autotuner
psdp_status_t psdp_autotune(psdp_handle_t h, int64_t n, const psdp_config_t* cfg) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  Plan p;
  STCHK(make_plan(cfg, n, p));
  void* base = nullptr;
  STCHK(arena_reserve(h, p.total, &base));
  Bufs b = carve(p, (char*)base);
  hipStream_t s = h->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  HIPCHK(launch_fill_sym(b.a32, p.ld, p.n, p.npad, h->seed, s));
  HIPCHK(launch_set_scale(h->st, 1.0, s));
  // every (format, averaged) pair the schedule uses
  std::vector<std::pair<int, int>> kinds;
  for (int i = 0; i < p.sch.nsteps; ++i) {
    kinds.emplace_back(p.sch.fmt[i], 0);
    kinds.emplace_back(p.sch.fmt[i], p.sch.avg[i]);
  }
  kinds.emplace_back(p.sch.final_fmt, p.sch.final_avg);
  std::sort(kinds.begin(), kinds.end());
  kinds.erase(std::unique(kinds.begin(), kinds.end()), kinds.end());
  for (auto [fmt, avg] : kinds) {
    HIPCHK(launch_scale_split(b.a32, p.N, p.ld, &h->st->inv_scale, fmt == FMT_F32 ? FMT_F32 : fmt, b.ap,
                              p.sch.final_fmt, b.x32, b.xp, s));
    ProdDesc d = base_desc(p);
    set_operand(d.x, fmt, b.a32, b.ap);
    set_operand(d.y, fmt, b.a32, b.ap);
    d.mode = EPI_STORE;
    d.c32 = fmt == FMT_F32 ? b.p3 : nullptr;
    d.out_fmt = fmt;
    for (int q = 0; q < 3; ++q) d.outp[q] = b.p2p[q];
    DESCCHK(d, fmt);
    const int top = max_sub(p.n, fmt, avg);
    int best = launch_sub(p.n, fmt, avg, h->cus);
    float best_ms = 1e30f;
    for (int sub = top; sub >= 2; sub /= 2) {
      STCHK(launch_product(h, p, fmt, avg, d, sub, s));  // warm-up (tile lists, caches)
      constexpr int R = 3;
      HIPCHK(hipEventRecord(e0, s));
      for (int r = 0; r < R; ++r) STCHK(launch_product(h, p, fmt, avg, d, sub, s));
      HIPCHK(hipEventRecord(e1, s));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) best_ms = ms, best = sub;
    }
    h->tuned[std::make_tuple(p.n, fmt, avg)] = best;
  }
  return PSDP_OK;
}

psdp_status_t psdp_set_tuned_tile(psdp_handle_t h, int64_t n, int format, int average, int tile) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (tile != 0 && tile != 64 && tile != 128 && tile != 256)
    return fail(PSDP_ERR_INVALID_ARG, "tile must be 64, 128, 256 (or 0 to clear)");
  std::lock_guard<std::mutex> lock(h->mu);
  const auto key = std::make_tuple((int)n, format, average ? 1 : 0);
  if (tile == 0) h->tuned.erase(key);
  else h->tuned[key] = tile / 32;
  return PSDP_OK;
}

psdp_status_t psdp_tuned_tile(psdp_handle_t h, int64_t n, int format, int average, int* tile) {
  if (!h || !tile) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lock(h->mu);
  auto it = h->tuned.find(std::make_tuple((int)n, format, average));
  *tile = it == h->tuned.end() ? 0 : 32 * it->second;
  return PSDP_OK;
}

psdp_status_t psdp_workspace_size(const psdp_config_t* cfg, int64_t n, size_t* bytes) {
  if (!bytes) return fail(PSDP_ERR_INVALID_ARG, "null output");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  Plan p;
  STCHK(make_plan(cfg, n, p));
  *bytes = p.total;
  return PSDP_OK;
}

psdp_status_t psdp_last_bound(psdp_handle_t h, double* lo, double* up, int* iterations) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  LanczosState st;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&st, h->st, sizeof st, hipMemcpyDeviceToHost));
  if (lo) *lo = st.lo;
  if (up) *up = st.up;
  if (iterations) *iterations = st.nb;
  return PSDP_OK;
}

psdp_status_t psdp_last_bounds_batched(psdp_handle_t h, double* lo, double* up, int* iterations, int count) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  if (!h->batch_states || count != h->batch_count)
    return fail(PSDP_ERR_INVALID_ARG, "the last batched call had %d blocks (asked for %d)", h->batch_count, count);
  std::vector<LanczosState> st((size_t)count);
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(st.data(), h->batch_states, sizeof(LanczosState) * count, hipMemcpyDeviceToHost));
  for (int b = 0; b < count; ++b) {
    if (lo) lo[b] = st[b].lo;
    if (up) up[b] = st[b].up;
    if (iterations) iterations[b] = st[b].nb;
  }
  return PSDP_OK;
}

psdp_status_t psdp_project_f64(psdp_handle_t h, const double* X, double* Y, int64_t n, const psdp_config_t* cfg,
                               void* ws, size_t ws_bytes) {
  return project(h, X, 1, Y, 1, n, cfg, ws, ws_bytes);
}

psdp_status_t psdp_project_f32(psdp_handle_t h, const float* X, float* Y, int64_t n, const psdp_config_t* cfg,
                               void* ws, size_t ws_bytes) {
  return project(h, X, 0, Y, 0, n, cfg, ws, ws_bytes);
}

// psdp_project_f32 without any device allocation: Lanczos state, tile lists
// and workspace come from the device's scratch array (layout: state, then
// the tile lists, then the workspace of the plan), whose address the handle
// looked up at creation.  Same kernels, tile sizes and order of work as
// psdp_project_f32 (tile lists in column order instead of the XCD-aware
// order, which changes no result).  Nothing here needs to start at zero:
// every buffer is written before it is read, as in the handle's arena.
static constexpr size_t NOALLOC_HEAD = 64 * 1024;  // state + tile lists
psdp_status_t psdp_project_f32_noalloc(psdp_handle_t h, const float* X, float* Y, int64_t n,
                                       const psdp_config_t* cfg) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (!X || !Y) return fail(PSDP_ERR_INVALID_ARG, "null matrix pointer");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  if (!h->scratch) return fail(PSDP_ERR_DEVICE, "the handle has no scratch array");
  Plan p;
  STCHK(make_plan(cfg, n, p));
  const size_t room = h->scratch_bytes - NOALLOC_HEAD;
  if (p.total > room)
    return fail(PSDP_ERR_INVALID_ARG,
                "n = %lld needs %zu bytes of workspace; the allocation-free entry point has %zu (use "
                "psdp_project_f32 for this size)",
                (long long)n, p.total, room);
  char* base = (char*)h->scratch;
  LanczosState* st = (LanczosState*)base;
  static_assert(sizeof(LanczosState) <= 4096, "state fits its slot");
  // one tile list per tile size the plan's products use
  FixedTiles ft;
  TileRef* lists = (TileRef*)(base + 4096);
  size_t used = 0;
  auto need = [&](int fmt, int avg) -> psdp_status_t {
    const int sub = tuned_sub(h, p, fmt, avg);
    if (sub != 2 && sub != 4)
      return fail(PSDP_ERR_UNSUPPORTED, "tile size %d is not available to the allocation-free entry point", 32 * sub);
    if (ft.list[sub]) return PSDP_OK;
    const int T = p.npad / (32 * sub);
    const size_t cnt = (size_t)T * (T + 1) / 2;
    if (4096 + (used + cnt) * sizeof(TileRef) > NOALLOC_HEAD)
      return fail(PSDP_ERR_INVALID_ARG, "internal: tile lists exceed their area");
    HIPCHK(launch_fill_tile_list(lists + used, T, h->stream));
    ft.list[sub] = lists + used;
    ft.count[sub] = (int)cnt;
    used += cnt;
    return PSDP_OK;
  };
  for (int i = 0; i < p.sch.nsteps; ++i) {
    STCHK(need(p.sch.fmt[i], 0));
    STCHK(need(p.sch.fmt[i], p.sch.avg[i]));
  }
  STCHK(need(p.sch.final_fmt, p.sch.final_avg));
  Bufs b = carve(p, base + NOALLOC_HEAD);
  const bool prof = h->profiling;
  h->profiling = false;  // timing events are created per launch
  const psdp_status_t est = enqueue_projection(h, p, b, X, 0, Y, 0, cfg, 1.0, st, &ft);
  h->profiling = prof;
  STCHK(est);
  if (cfg->synchronize) STCHK(check_flags(h, st));
  return PSDP_OK;
}

static psdp_status_t dropin(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes, int method,
                            int auto_scale) {
  psdp_config_t cfg;
  psdp_default_config(&cfg, method);
  cfg.auto_scale = auto_scale;
  return project(h, mat, 1, mat, 1, n, &cfg, ws, ws_bytes);
}

psdp_status_t psdp_composite_fp32(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP32, 0);
}
psdp_status_t psdp_composite_fp32_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                             size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP32, 1);
}
psdp_status_t psdp_composite_fp16(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP16, 0);
}
psdp_status_t psdp_composite_fp16_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                             size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP16, 1);
}
psdp_status_t psdp_composite_fp32_emulated(psdp_handle_t h, double* mat, int64_t n, void* ws, size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP32_EMULATED, 0);
}
psdp_status_t psdp_composite_fp32_emulated_auto_scale(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                                      size_t ws_bytes) {
  return dropin(h, mat, n, ws, ws_bytes, PSDP_METHOD_COMPOSITE_FP32_EMULATED, 1);
}

// approximate_two_norm on a float64 matrix; caller holds h->mu.
static psdp_status_t two_norm_locked(psdp_handle_t h, const double* A, int64_t n, int64_t max_iter, double tol,
                                     double* lo, double* up) {
  const int iters = (int)std::min<int64_t>(max_iter, n);
  const size_t vb = align256((size_t)n * 8);
  void* base = nullptr;
  STCHK(arena_reserve(h, vb * 6 + align256((size_t)n * 8 * iters), &base));
  char* v = (char*)base;
  LanczosVecs lv;
  lv.q = (double*)(v + 0 * vb);
  lv.w = (double*)(v + 1 * vb);
  lv.w1 = (double*)(v + 2 * vb);
  lv.vold = (double*)(v + 3 * vb);
  lv.y = (double*)(v + 4 * vb);
  lv.ay = (double*)(v + 5 * vb);
  lv.V = (double*)(v + 6 * vb);
  lv.part = nullptr;  // float64 matrix: full-read GEMV
  HIPCHK(launch_lanczos(A, 1, n, (int)n, (int)n, iters, tol, h->seed, 1.0, 0, h->st, lv, h->stream));
  LanczosState st;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&st, h->st, sizeof st, hipMemcpyDeviceToHost));
  *lo = st.lo;
  *up = st.up;
  if (st.nonfinite) return fail(PSDP_ERR_NONFINITE, "non-finite value met by the Lanczos bound");
  return PSDP_OK;
}

psdp_status_t psdp_approximate_two_norm(psdp_handle_t h, const double* A, int64_t n, double* lo, double* up,
                                        int64_t max_iter, double tol) {
  if (!h || !A || !lo || !up) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0) return fail(PSDP_ERR_INVALID_ARG, "matrix size must be positive");
  if (max_iter <= 0) max_iter = 20;
  if (max_iter > LANCZOS_MAX_ITER)
    return fail(PSDP_ERR_INVALID_ARG, "at most %d Lanczos iterations are supported", LANCZOS_MAX_ITER);
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  ProfScope ps(h, PSDP_PROF_LANCZOS);
  return two_norm_locked(h, A, n, max_iter, tol, lo, up);
}

psdp_status_t psdp_sym_product_f32(psdp_handle_t h, const float* X, const float* Y, float* C, int64_t n,
                                   int format, int average) {
  if (!h || !X || !Y || !C) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0 || n > (1 << 20)) return fail(PSDP_ERR_INVALID_ARG, "bad matrix size");
  if (format < 0 || format >= FMT_COUNT) return fail(PSDP_ERR_INVALID_ARG, "unknown format %d", format);
  if (average && !avg_capable(format))
    return fail(PSDP_ERR_UNSUPPORTED, "averaged products exist for binary16 and bfloat16 operands only");
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  // the projections' tile choice for this size (tuned or built in); n is
  // padded for the largest tile, as in a projection plan
  const int avg = average ? 1 : 0;
  const int top = max_sub(n, format, avg);
  auto it = h->tuned.find(std::make_tuple((int)n, format, avg));
  const int sub = it != h->tuned.end() && it->second <= top ? it->second : launch_sub(n, format, avg, h->cus);
  const int bt = 32 * sub, btop = 32 * top;
  const int npad = (int)((n + btop - 1) / btop * btop), ld = npad + LD_PAD;
  const int64_t N = (int64_t)npad * ld;
  const size_t f32b = align256((size_t)N * 4), u16b = align256((size_t)N * 2);
  void* base = nullptr;
  STCHK(arena_reserve(h, 2 * f32b + 6 * u16b, &base));
  char* c = (char*)base;
  float* mx = (float*)c;
  float* my = (float*)(c + f32b);
  uint16_t* px[3];
  uint16_t* py[3];
  for (int i = 0; i < 3; ++i) {
    px[i] = (uint16_t*)(c + 2 * f32b + i * u16b);
    py[i] = (uint16_t*)(c + 2 * f32b + (3 + i) * u16b);
  }
  hipStream_t s = h->stream;
  TileList tl;
  if (sub != 8) STCHK(get_tiles(h, npad / bt, bt, tl));
  {
    ProfScope ps(h, PSDP_PROF_PREP);
    HIPCHK(launch_sym_pad(X, 0, n, (int)n, mx, ld, npad, s));
    HIPCHK(launch_sym_pad(Y, 0, n, (int)n, my, ld, npad, s));
    uint16_t* none[3] = {nullptr, nullptr, nullptr};
    HIPCHK(launch_scale_split(mx, N, ld, nullptr, format, px, FMT_F32, mx, none, s));
    HIPCHK(launch_scale_split(my, N, ld, nullptr, format, py, FMT_F32, my, none, s));
  }
  ProdDesc d;
  memset(&d, 0, sizeof d);
  d.ld = ld;
  d.npad = npad;
  d.n = (int)n;
  set_operand(d.x, format, mx, px);
  set_operand(d.y, format, my, py);
  d.mode = EPI_OUTPUT;
  d.out = C;
  d.out_ld = n;
  d.out_f64 = 0;
  d.scale = nullptr;
  {
    ProfScope ps(h, PSDP_PROF_PROD_F32 + format);
    if (sub == 8) STCHK(launch_big_product(h, npad / bt, format, average ? 1 : 0, d, s));  // as the projections
    else HIPCHK(launch_sym_product(format, sub, average ? 1 : 0, d, tl.dev, tl.count, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return PSDP_OK;
}

// Many independent blocks (a block-diagonal cone): every block keeps its own
// bound, scale and buffers; each stage of the chain is ONE persistent grouped
// launch over the upper-triangle tiles of all blocks (64 x 64 tiles, sorted by
// K length so that the static round-robin over workgroups stays balanced).
// A block's result is bitwise the same alone and in any batch: the same tile
// code, k order and bound kernels run on it either way.
// Blocks with n >= BIG_BLOCK_N use 128 x 128 tiles (their own grouped launch
// per stage), smaller ones 64 x 64: a block's tile size depends on its own n
// only, so its result does not depend on the batch.  Measured on 256 blocks
// of n in [32, 1024]: moving the n >= 512 blocks to 128 x 128 tiles made the
// batch 9% slower (short K loops, two launches in series), hence 2048.
static constexpr int BIG_BLOCK_N = 2048;

// Tile list of one class of a grouped launch.  Workgroups are dispatched
// round-robin over the 8 XCDs and the persistent grid (a multiple of 8, or
// one workgroup per entry) keeps that residue, so list position p runs on XCD
// p % 8.  Blocks are spread over the XCDs by cost (T(T+1)/2 tiles of K =
// npad, largest first onto the least loaded XCD) so that all tiles of a
// block share one XCD's L2; an XCD's tiles are grouped by block with K
// decreasing (the workgroups' static round-robin over them stays balanced);
// the 8 lists are interleaved and the short ones padded with empty entries
// (mat = -1).
static std::vector<TileRef> batched_tile_order(const std::vector<int>& blocks, int bt,
                                               const std::vector<int>& npad) {
  auto npad_of = [&](int b) { return npad[b]; };
  constexpr int XCD = 8;
  std::vector<int> order(blocks);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return npad_of(a) > npad_of(b); });
  std::vector<std::vector<TileRef>> lists(XCD);
  double load[XCD] = {};
  for (int b : order) {
    const int T = npad_of(b) / bt;
    const double cost = 0.5 * T * (T + 1) * (double)npad_of(b);
    int x = 0;
    for (int k = 1; k < XCD; ++k)
      if (load[k] < load[x]) x = k;
    load[x] += cost;
    for (int i = 0; i < T; ++i)
      for (int j = i; j < T; ++j) lists[x].push_back(TileRef{b, i, j, 0});
  }
  size_t len = 0;
  for (auto& l : lists) len = std::max(len, l.size());
  std::vector<TileRef> out;
  out.reserve(len * XCD);
  for (size_t k = 0; k < len; ++k)
    for (int x = 0; x < XCD; ++x) out.push_back(k < lists[x].size() ? lists[x][k] : TileRef{-1, 0, 0, 0});
  // trailing padding entries are not needed
  while (!out.empty() && out.back().mat < 0) out.pop_back();
  return out;
}

// Buffers of one block of a grouped projection, as byte offsets from the
// workspace base (kNone: a buffer the schedule does not use).  Blocks are
// small: rows are not padded beyond the tile (ld = npad).
struct BlockBufs {
  int n = 0, npad = 0, iters = 0;
  size_t a32 = 0, p3 = 0, p2 = kNone, x32 = kNone, vec = 0;
  size_t ap[3] = {kNone, kNone, kNone}, p2p[3] = {kNone, kNone, kNone};
  size_t p3p[3] = {kNone, kNone, kNone}, xp[3] = {kNone, kNone, kNone};
};

// Tile edge of a block's grouped products: 128 from big_n, else 64 (a
// function of the block's own size only, so its result never depends on
// the rest of the batch).
static int block_tile(int n, int big_n) { return n >= big_n ? 128 : 64; }

// Lays out one block's buffers from byte offset `off`; returns the end.
static size_t layout_block(const Schedule& sch, int n, int iters, int big_n, size_t off, BlockBufs& q) {
  bool any_f32 = false;
  int pa = pieces_of(sch.final_fmt), pp = 0;
  for (int i = 0; i < sch.nsteps; ++i) {
    any_f32 |= sch.fmt[i] == FMT_F32;
    pa = std::max(pa, pieces_of(sch.fmt[i]));
    pp = std::max(pp, pieces_of(sch.fmt[i]));
  }
  const int px = pieces_of(sch.final_fmt);
  q = BlockBufs();
  q.n = n;
  const int bt = block_tile(n, big_n);
  q.npad = (n + bt - 1) / bt * bt;
  q.iters = std::min(iters, n);
  const int64_t N = (int64_t)q.npad * q.npad;
  const size_t f32b = align256((size_t)N * 4);
  q.a32 = off, off += f32b;
  q.p3 = off, off += f32b;
  if (any_f32) q.p2 = off, off += f32b;
  if (sch.final_fmt == FMT_F32) q.x32 = off, off += f32b;
  size_t ab[3], sb[3];
  piece_sizes(sch, q.npad, q.npad, ab, sb);
  for (int i = 0; i < pa; ++i) q.ap[i] = off, off += ab[i];
  for (int i = 0; i < pp; ++i) q.p2p[i] = off, off += sb[i];
  for (int i = 0; i < pp; ++i) q.p3p[i] = off, off += sb[i];
  for (int i = 0; i < px; ++i) q.xp[i] = off, off += piece_bytes(sch.final_fmt, i, q.npad, q.npad);
  q.vec = off;
  off += align256((size_t)q.npad * 8) * 6 + align256((size_t)q.npad * 8 * q.iters);
  return off;
}

// Launch descriptors of one block laid out by layout_block at `base`: its
// product descriptor of stage s at pd[s * stride] (3 per filter step, then
// the output product), its Lanczos, pad and scale/split jobs.  st: the
// block's device Lanczos state.  Pure host code.
static void block_descs(const Schedule& sch, const BlockBufs& q, char* base, const float* X, float* Y,
                        LanczosState* st, ProdDesc* pd, size_t stride, LanczosJob& J, PadJob& pj, SplitJob& S) {
  auto F = [&](size_t o) { return o != kNone ? (float*)(base + o) : (float*)nullptr; };
  auto U = [&](size_t o) { return o != kNone ? (uint16_t*)(base + o) : (uint16_t*)nullptr; };
  Bufs B;
  B.a32 = F(q.a32);
  B.p3 = F(q.p3);
  B.p2 = F(q.p2);
  B.x32 = F(q.x32);
  for (int i = 0; i < 3; ++i) {
    B.ap[i] = U(q.ap[i]);
    B.p2p[i] = U(q.p2p[i]);
    B.p3p[i] = U(q.p3p[i]);
    B.xp[i] = U(q.xp[i]);
  }
  const size_t vb = align256((size_t)q.npad * 8);
  char* v = base + q.vec;
  J.A = B.a32;
  J.ld = q.npad;
  J.n = q.n;
  J.m = q.npad;
  J.max_iter = q.iters;
  J.st = st;
  J.v.q = (double*)(v + 0 * vb);
  J.v.w = (double*)(v + 1 * vb);
  J.v.w1 = (double*)(v + 2 * vb);
  J.v.vold = (double*)(v + 3 * vb);
  J.v.y = (double*)(v + 4 * vb);
  J.v.ay = (double*)(v + 5 * vb);
  J.v.V = (double*)(v + 6 * vb);
  J.v.part = (double*)B.p3;  // GEMV partial sums; A^3 is written only after the bound
  pj = PadJob{X, q.n, q.npad, B.a32};
  S.a32 = B.a32;
  S.count = (int64_t)q.npad * q.npad;
  S.ld = q.npad;
  S.inv = &st->inv_scale;
  S.x32 = B.x32;
  for (int i = 0; i < 3; ++i) {
    S.ap[i] = B.ap[i];
    S.xp[i] = B.xp[i];
  }
  Plan p;
  p.n = q.n;
  p.npad = q.npad;
  p.ld = q.npad;
  for (int i = 0; i < sch.nsteps; ++i) {
    const int f = sch.fmt[i];
    const int fn = i + 1 < sch.nsteps ? sch.fmt[i + 1] : sch.final_fmt;
    const bool last = i + 1 == sch.nsteps;
    ProdDesc& d1 = pd[(size_t)(3 * i) * stride];
    d1 = base_desc(p);
    set_operand(d1.x, f, B.a32, B.ap);
    set_operand(d1.y, f, B.a32, B.ap);
    d1.mode = EPI_STORE;
    d1.c32 = f == FMT_F32 ? B.p2 : nullptr;
    d1.out_fmt = f;
    for (int k = 0; k < 3; ++k) d1.outp[k] = B.p2p[k];
    ProdDesc& d2 = pd[(size_t)(3 * i + 1) * stride];
    d2 = base_desc(p);
    set_operand(d2.x, f, B.p2, B.p2p);
    set_operand(d2.y, f, B.a32, B.ap);
    d2.mode = EPI_STORE;
    d2.c32 = a3_f32(f) ? B.p3 : nullptr;
    d2.out_fmt = f;
    for (int k = 0; k < 3; ++k) d2.outp[k] = B.p3p[k];
    ProdDesc& d3 = pd[(size_t)(3 * i + 2) * stride];
    d3 = base_desc(p);
    set_operand(d3.x, f, B.p3, B.p3p);
    set_operand(d3.y, f, B.p2, B.p2p);
    d3.mode = last ? EPI_POLY_HALF_ID : EPI_POLY;
    d3.a = sch.a[i];
    d3.b = sch.b[i];
    d3.c = sch.c[i];
    d3.c32 = B.a32;
    d3.p3 = a3_f32(f) ? B.p3 : nullptr;
    d3.out_fmt = fn;
    for (int k = 0; k < 3; ++k) d3.outp[k] = B.ap[k];
  }
  ProdDesc& d4 = pd[(size_t)(3 * sch.nsteps) * stride];
  d4 = base_desc(p);
  set_operand(d4.x, sch.final_fmt, B.x32, B.xp);
  set_operand(d4.y, sch.final_fmt, B.a32, B.ap);
  d4.mode = EPI_OUTPUT;
  d4.out = Y;
  d4.out_ld = q.n;
  d4.out_f64 = 0;
  d4.scale = &st->scale;
  d4.nonfinite = &st->out_nonfinite;
}

struct BatchPlan {
  Schedule sch;
  int iters = 20, tight = 0;
  double tol = 1e-10, margin = 1.0;
  int count = 0, nstages = 0, max_npad = 0;
  int ntiles_big = 0, ntiles_small = 0;  // 128 x 128 tiles first, then 64 x 64
  int npad_tiles = 0, nsymv_tiles = 0, nsymv_chunks = 0, nsplit_chunks = 0;  // flat work lists
  int64_t max_N = 0;
  size_t off_states = 0, off_desc = 0, total = 0;
  size_t d_states = 0, d_lz = 0, d_pad = 0, d_split = 0, d_prod = 0, d_tiles = 0;
  size_t d_padt = 0, d_symt = 0, d_rows = 0, d_chunks = 0;
  std::vector<char> blob;  // initial states + descriptors, built for one workspace base
};

// Lays out the workspace and, when base is given, builds every descriptor for
// that base and checks them.  Pure host code: no device call.
static psdp_status_t plan_batched(const psdp_block_t* blocks, int count, const psdp_config_t* cfg, char* base,
                                  BatchPlan& bp, int big_n = BIG_BLOCK_N) {
  if (!blocks) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (count <= 0 || count > 65535) return fail(PSDP_ERR_INVALID_ARG, "block count must be in [1, 65535]");
  Schedule& sch = bp.sch;
  STCHK(build_schedule(cfg, sch));
  bp.iters = cfg->lanczos_iters > 0 ? cfg->lanczos_iters : 20;
  if (bp.iters > LANCZOS_MAX_ITER)
    return fail(PSDP_ERR_INVALID_ARG, "at most %d Lanczos iterations are supported", LANCZOS_MAX_ITER);
  bp.tol = cfg->lanczos_tol > 0 ? cfg->lanczos_tol : 1e-10;
  bp.margin = cfg->scale_margin > 0 ? cfg->scale_margin : 1.0;
  bp.tight = cfg->bound_mode == PSDP_BOUND_TIGHT;
  for (int b = 0; b < count; ++b)
    if (!blocks[b].X || !blocks[b].Y || blocks[b].n <= 0 || blocks[b].n > 65536)
      return fail(PSDP_ERR_INVALID_ARG, "block %d: bad pointer or size", b);
  bp.count = count;
  std::vector<BlockBufs> bb(count);
  size_t off = 0;
  for (int b = 0; b < count; ++b) {
    off = layout_block(sch, (int)blocks[b].n, bp.iters, big_n, off, bb[b]);
    bp.max_npad = std::max(bp.max_npad, bb[b].npad);
    bp.max_N = std::max(bp.max_N, (int64_t)bb[b].npad * bb[b].npad);
  }
  // tiles of each class in XCD-aware order (batched_tile_order)
  std::vector<int> big_blocks, small_blocks;
  for (int b = 0; b < count; ++b) (bb[b].n >= big_n ? big_blocks : small_blocks).push_back(b);
  std::vector<int> npads(count);
  for (int b = 0; b < count; ++b) npads[b] = bb[b].npad;
  std::vector<TileRef> tiles = batched_tile_order(big_blocks, 128, npads);
  std::vector<TileRef> small = batched_tile_order(small_blocks, 64, npads);
  bp.ntiles_big = (int)tiles.size();
  bp.ntiles_small = (int)small.size();
  tiles.insert(tiles.end(), small.begin(), small.end());
  // flat work lists: 64 x 64 tile pairs of the symmetrize/pad, upper-triangle
  // tiles and row chunks of the Lanczos GEMV, value chunks of the scale/split
  std::vector<TileRef> padt, symt;
  std::vector<int2> rows, chunks;
  for (int b = 0; b < count; ++b) {
    const int T = bb[b].npad / 64;
    for (int i = 0; i < T; ++i)
      for (int j = i; j < T; ++j) padt.push_back(TileRef{b, i, j, 0});
    const int nt = (bb[b].npad + SYMV_TILE - 1) / SYMV_TILE;
    for (int j = 0; j < nt; ++j)
      for (int i = 0; i <= j; ++i) symt.push_back(TileRef{b, i, j, 0});
    for (int r = 0; r < nt; ++r) rows.push_back(make_int2(b, r));
    const int64_t nc = (int64_t)bb[b].npad * bb[b].npad / SPLIT_CHUNK;
    for (int64_t c = 0; c < nc; ++c) chunks.push_back(make_int2(b, (int)c));
  }
  bp.npad_tiles = (int)padt.size();
  bp.nsymv_tiles = (int)symt.size();
  bp.nsymv_chunks = (int)rows.size();
  bp.nsplit_chunks = (int)chunks.size();
  bp.nstages = 3 * sch.nsteps + 1;
  bp.off_states = off;
  off += align256(sizeof(LanczosState) * count);
  bp.off_desc = off;  // everything below is uploaded in one copy
  size_t doff = 0;
  bp.d_states = doff, doff += align256(sizeof(LanczosState) * count);  // initial states (copied to off_states)
  bp.d_lz = doff, doff += align256(sizeof(LanczosJob) * count);
  bp.d_pad = doff, doff += align256(sizeof(PadJob) * count);
  bp.d_split = doff, doff += align256(sizeof(SplitJob) * count);
  bp.d_prod = doff, doff += align256(sizeof(ProdDesc) * (size_t)bp.nstages * count);
  bp.d_tiles = doff, doff += align256(sizeof(TileRef) * tiles.size());
  bp.d_padt = doff, doff += align256(sizeof(TileRef) * padt.size());
  bp.d_symt = doff, doff += align256(sizeof(TileRef) * symt.size());
  bp.d_rows = doff, doff += align256(sizeof(int2) * rows.size());
  bp.d_chunks = doff, doff += align256(sizeof(int2) * chunks.size());
  bp.total = off + doff;
  if (!base) return PSDP_OK;

  LanczosState* states = (LanczosState*)(base + bp.off_states);
  bp.blob.assign(doff, 0);
  char* blob = bp.blob.data();
  LanczosState* st0 = (LanczosState*)(blob + bp.d_states);  // scale 1 (used as is without auto_scale)
  for (int b = 0; b < count; ++b) st0[b].scale = st0[b].inv_scale = st0[b].up = 1.0;
  LanczosJob* lz = (LanczosJob*)(blob + bp.d_lz);
  PadJob* pj = (PadJob*)(blob + bp.d_pad);
  SplitJob* sj = (SplitJob*)(blob + bp.d_split);
  ProdDesc* pd = (ProdDesc*)(blob + bp.d_prod);
  memcpy(blob + bp.d_tiles, tiles.data(), sizeof(TileRef) * tiles.size());
  memcpy(blob + bp.d_padt, padt.data(), sizeof(TileRef) * padt.size());
  memcpy(blob + bp.d_symt, symt.data(), sizeof(TileRef) * symt.size());
  memcpy(blob + bp.d_rows, rows.data(), sizeof(int2) * rows.size());
  memcpy(blob + bp.d_chunks, chunks.data(), sizeof(int2) * chunks.size());

  for (int b = 0; b < count; ++b)
    block_descs(sch, bb[b], base, blocks[b].X, blocks[b].Y, states + b, pd + b, (size_t)count, lz[b], pj[b], sj[b]);
  for (int st = 0; st < bp.nstages; ++st) {
    const int f = st >= 3 * sch.nsteps ? sch.final_fmt : sch.fmt[st / 3];
    for (int b = 0; b < count; ++b) DESCCHK(pd[(size_t)st * count + b], f);
  }
  for (int b = 0; b < count; ++b) {
    if (!pj[b].X || !pj[b].M || !sj[b].a32 || !lz[b].A || !lz[b].v.q || !lz[b].v.V)
      return fail(PSDP_ERR_INVALID_ARG, "internal: block %d has an unset buffer", b);
    if (sch.fmt[0] != FMT_F32 && !sj[b].ap[0]) return fail(PSDP_ERR_INVALID_ARG, "internal: block %d: no A pieces", b);
    if (sch.final_fmt == FMT_F32 ? !sj[b].x32 : !sj[b].xp[0])
      return fail(PSDP_ERR_INVALID_ARG, "internal: block %d: no X_s buffer", b);
  }
  return PSDP_OK;
}

psdp_status_t psdp_workspace_size_batched(const psdp_block_t* blocks, int count, const psdp_config_t* cfg,
                                          size_t* bytes) {
  if (!bytes) return fail(PSDP_ERR_INVALID_ARG, "null output");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  BatchPlan bp;
  STCHK(plan_batched(blocks, count, cfg, nullptr, bp));
  // Build and check every descriptor against a stand-in base address: the
  // same host code as a real call, without touching a device.
  BatchPlan chk;
  STCHK(plan_batched(blocks, count, cfg, (char*)(uintptr_t)0x10000000, chk));
  *bytes = bp.total;
  return PSDP_OK;
}

psdp_status_t psdp_project_batched_f32(psdp_handle_t h, const psdp_block_t* blocks, int count,
                                       const psdp_config_t* cfg) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  std::lock_guard<std::mutex> lock(h->mu);
  const int big_n = h->batch_big_n > 0 ? h->batch_big_n : BIG_BLOCK_N;
  BatchPlan bp;
  STCHK(plan_batched(blocks, count, cfg, nullptr, bp, big_n));
  HIPCHK(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  void* base_v = nullptr;
  STCHK(arena_reserve(h, bp.total, &base_v));
  char* base = (char*)base_v;
  STCHK(plan_batched(blocks, count, cfg, base, bp, big_n));
  const bool two = h->batch_concurrent && bp.ntiles_big && bp.ntiles_small;
  if (two && !h->aux) {
    HIPCHK(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->join_ev, hipEventDisableTiming));
  }
  const Schedule& sch = bp.sch;
  LanczosState* states = (LanczosState*)(base + bp.off_states);
  char* ddev = base + bp.off_desc;

  // descriptors and initial states go up from a pinned staging slot; the
  // host waits only if that slot's previous copy has not run yet
  char* stage = nullptr;
  STCHK(staging_slot(h, bp.blob.size(), &stage));
  memcpy(stage, bp.blob.data(), bp.blob.size());
  HIPCHK(hipMemcpyAsync(ddev, stage, bp.blob.size(), hipMemcpyHostToDevice, s));
  STCHK(staging_release(h));

  h->batch_states = states;
  h->batch_count = count;
  const LanczosJob* lz_d = (const LanczosJob*)(ddev + bp.d_lz);
  const ProdDesc* pd_d = (const ProdDesc*)(ddev + bp.d_prod);
  const TileRef* tl_d = (const TileRef*)(ddev + bp.d_tiles);
  static_assert(sizeof(LanczosState) % 8 == 0, "states are copied in 8-byte words");
  // everything after the upload reads only device memory: one fixed sequence
  // of launches, which graph mode captures once and replays
  auto enqueue = [&](hipStream_t s) -> psdp_status_t {
    HIPCHK(launch_copy_words(states, ddev + bp.d_states, (int64_t)(sizeof(LanczosState) * count / 8), s));
    {
      ProfScope ps(h, PSDP_PROF_PREP);
      HIPCHK(launch_sym_pad_batched((const PadJob*)(ddev + bp.d_pad), (const TileRef*)(ddev + bp.d_padt),
                                    bp.npad_tiles, s));
    }
    if (cfg->auto_scale) {
      ProfScope ps(h, PSDP_PROF_LANCZOS);
      HIPCHK(launch_lanczos_batched(lz_d, count, bp.max_npad, bp.iters, bp.tol, h->seed, bp.margin, bp.tight,
                                    (const TileRef*)(ddev + bp.d_symt), bp.nsymv_tiles,
                                    (const int2*)(ddev + bp.d_rows), bp.nsymv_chunks, s));
    }
    {
      ProfScope ps(h, PSDP_PROF_PREP);
      HIPCHK(launch_scale_split_batched((const SplitJob*)(ddev + bp.d_split), (const int2*)(ddev + bp.d_chunks),
                                        bp.nsplit_chunks, sch.fmt[0], sch.final_fmt, s));
    }
    for (int st = 0; st < bp.nstages; ++st) {
      const bool fin = st >= 3 * sch.nsteps;
      const int f = fin ? sch.final_fmt : sch.fmt[st / 3];
      const int avg = fin ? sch.final_avg : (st % 3 != 0 ? sch.avg[st / 3] : 0);
      ProfScope ps(h, PSDP_PROF_PROD_F32 + f);
      const ProdDesc* pd = pd_d + (size_t)st * count;
      // the small-tile class runs beside the big one on the second stream
      hipStream_t s2 = two ? h->aux : s;
      if (two) {
        HIPCHK(hipEventRecord(h->fork_ev, s));
        HIPCHK(hipStreamWaitEvent(h->aux, h->fork_ev, 0));
      }
      if (bp.ntiles_big)
        HIPCHK(launch_sym_product_grouped(f, 4, avg, pd, tl_d, bp.ntiles_big, std::min(bp.ntiles_big, h->cus), s));
      // persistent grids of exactly the resident workgroups (two 64 x 64 tiles
      // per CU): workgroup w then runs list positions w, w + 2 cus, ..., so the
      // tiles an XCD runs together are neighbours in its block-ordered list (a
      // grid of 4 cus, half of it waiting to start, spread them over the list:
      // 7.63 against 7.45-7.48 ms per mixed_mx batch, bitwise equal)
      if (bp.ntiles_small)
        HIPCHK(launch_sym_product_grouped(f, 2, avg, pd, tl_d + bp.ntiles_big, bp.ntiles_small,
                                          std::min(bp.ntiles_small, 2 * h->cus), s2));
      if (two) {
        HIPCHK(hipEventRecord(h->join_ev, h->aux));
        HIPCHK(hipStreamWaitEvent(s, h->join_ev, 0));
      }
    }
    return PSDP_OK;
  };
  if (!h->batch_graph) {
    STCHK(enqueue(s));
  } else {
    // the graph holds kernel arguments (workspace addresses, sizes, seed,
    // streams): replay it only for the same plan on the same workspace
    auto bits = [](double v) {
      int64_t b;
      memcpy(&b, &v, sizeof b);
      return b;
    };
    std::vector<int64_t> key = {count, big_n, two, (int64_t)(intptr_t)base, (int64_t)h->seed,
                                (int64_t)(intptr_t)h->aux, cfg->auto_scale, bp.iters,
                                bits(bp.tol), bits(bp.margin), sch.nsteps, sch.final_fmt, sch.final_avg};
    for (int i = 0; i < sch.nsteps; ++i)
      for (int64_t v : {(int64_t)sch.fmt[i], (int64_t)sch.avg[i], bits(sch.a[i]), bits(sch.b[i]), bits(sch.c[i])})
        key.push_back(v);
    for (int b = 0; b < count; ++b) key.push_back(blocks[b].n);
    if (!h->bgraph || key != h->bgraph_key) {
      if (h->bgraph) {
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipGraphExecDestroy(h->bgraph));
        h->bgraph = nullptr;
      }
      // captured on a stream of the handle's own (the caller's stream may be
      // the legacy default stream, which cannot capture); replayed on s
      if (!h->cap) HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
      const bool prof = h->profiling;
      h->profiling = false;  // no timing events inside the graph
      HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
      const psdp_status_t est = enqueue(h->cap);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(h->cap, &g);
      h->profiling = prof;
      STCHK(est);
      HIPCHK(ce);
      const hipError_t ie = hipGraphInstantiate(&h->bgraph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(ie);
      h->bgraph_key = key;
    }
    ProfScope ps(h, PSDP_PROF_GRAPH);
    HIPCHK(hipGraphLaunch(h->bgraph, s));
  }
  if (cfg->synchronize) {
    std::vector<LanczosState> out(count);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(out.data(), states, sizeof(LanczosState) * count, hipMemcpyDeviceToHost));
    for (int b = 0; b < count; ++b) {
      if (out[b].nonfinite) return fail(PSDP_ERR_NONFINITE, "block %d: non-finite value in the input", b);
      if (out[b].out_nonfinite)
        return fail(PSDP_ERR_NONFINITE, "block %d: non-finite value in the result (diverged chain)", b);
    }
  }
  return PSDP_OK;
}

// ---------------------------------------------------------------- session (continuous batching)
// A changing set of blocks advanced one product per tick.  Each block owns a
// region of the session's fixed arena (buffers as layout_block places them,
// then its Lanczos state) and its host-side launch descriptors
// (block_descs).  A block goes through
//   * its prologue -- symmetrize/pad, the batched Lanczos bound and the
//     scale/split, the kernels of the batched call -- run for all blocks
//     added since the last prologue on the handle's second stream, so that
//     the latency-bound Lanczos launches overlap the products;
//   * once the host has seen that prologue's event complete, one product
//     per tick: a tick is, on the handle's stream, one persistent grouped
//     launch per (operand format, averaged, tile class) over the next
//     product of every ready block -- the same kernels and descriptors as
//     the batched call, so a block's result is bitwise the one it gets alone.
// Descriptors and work lists go up in one copy per prologue / tick from the
// handle's pinned staging ring into the session's device areas (one per
// stream; stream order keeps each copy behind the kernels that read the
// previous one).  A block's region is reused only after the event of the
// tick that ran its output product has completed.
struct SessBlock {
  int64_t ticket = 0;
  size_t off = 0, bytes = 0;  // arena region
  BlockBufs q;
  LanczosState* st = nullptr;
  std::vector<ProdDesc> pd;   // one per stage
  LanczosJob lz;
  PadJob pj;
  SplitJob sj;
  int stage = -2;             // -2: prologue not enqueued, -1: prologue running, then the next product
};

struct DevArea {
  char* p = nullptr;
  size_t bytes = 0;
};

struct psdp_session_st {
  psdp_handle_t h = nullptr;
  Schedule sch;
  int iters = 20, tight = 0, big_n = BIG_BLOCK_N, nstages = 0;
  double tol = 1e-10, margin = 1.0;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  DevArea tick_area, pro_area;
  hipStream_t pro_stream = nullptr;  // prologues
  struct Prologue {
    hipEvent_t ev;
    std::vector<int64_t> tickets;
  };
  std::deque<Prologue> pros;         // enqueued, not yet seen complete (in stream order)
  std::map<size_t, size_t> free_;  // offset -> size
  std::vector<SessBlock> blocks;   // admitted, not finished
  struct Done {
    int64_t ticket;
    size_t off, bytes;
  };
  struct Tick {
    hipEvent_t ev;
    std::vector<Done> finished;
  };
  std::deque<Tick> ticks;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int64_t> completed;
  int64_t next_ticket = 1;
};

// Arena bytes of a block of size n: its buffers, then its Lanczos state.
static size_t session_block_bytes(const psdp_session_st* s, int n, size_t* state_off) {
  BlockBufs q;
  const size_t end = layout_block(s->sch, n, s->iters, s->big_n, 0, q);
  *state_off = align256(end);
  return *state_off + align256(sizeof(LanczosState));
}

// First fit over the free regions (offset -> size, coalesced on free).
static bool session_alloc(psdp_session_st* s, size_t bytes, size_t* off) {
  for (auto it = s->free_.begin(); it != s->free_.end(); ++it) {
    if (it->second < bytes) continue;
    *off = it->first;
    const size_t rest = it->second - bytes;
    const size_t at = it->first + bytes;
    s->free_.erase(it);
    if (rest) s->free_[at] = rest;
    return true;
  }
  return false;
}

static void session_free(psdp_session_st* s, size_t off, size_t bytes) {
  auto it = s->free_.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != s->free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    s->free_.erase(nx);
  }
  if (it != s->free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      s->free_.erase(it);
    }
  }
}

// true when the event has completed (false: not yet); errors become a status
static psdp_status_t event_done(hipEvent_t e, bool* done) {
  const hipError_t q = hipEventQuery(e);
  if (q == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = false;
    return PSDP_OK;
  }
  HIPCHK(q);
  *done = true;
  return PSDP_OK;
}

// Retires completed ticks, oldest first (their finished blocks give back
// their regions and report their tickets), and makes the blocks of a
// completed prologue ready for ticks.  Caller holds h->mu.
static psdp_status_t session_retire(psdp_session_st* s) {
  while (!s->ticks.empty()) {
    bool d = false;
    STCHK(event_done(s->ticks.front().ev, &d));
    if (!d) break;
    auto& t = s->ticks.front();
    for (const auto& f : t.finished) {
      session_free(s, f.off, f.bytes);
      s->completed.push_back(f.ticket);
    }
    s->ev_pool.push_back(t.ev);
    s->ticks.pop_front();
  }
  while (!s->pros.empty()) {
    bool d = false;
    STCHK(event_done(s->pros.front().ev, &d));
    if (!d) break;
    const auto& tk = s->pros.front().tickets;
    for (auto& b : s->blocks)
      if (b.stage == -1 && std::find(tk.begin(), tk.end(), b.ticket) != tk.end()) b.stage = 0;
    s->ev_pool.push_back(s->pros.front().ev);
    s->pros.pop_front();
  }
  return PSDP_OK;
}

extern "C++" {
// Appends a typed array to a host blob (256-byte aligned); returns its offset.
template <typename T>
static size_t blob_put(std::vector<char>& blob, const T* data, size_t count) {
  const size_t off = align256(blob.size());
  blob.resize(off + align256(sizeof(T) * std::max<size_t>(count, 1)), 0);
  if (count) memcpy(blob.data() + off, data, sizeof(T) * count);
  return off;
}
}

// Copies a host blob to a session device area on `st` (through the handle's
// staging ring); grows the area first if needed.
static psdp_status_t session_upload(psdp_session_st* s, DevArea& a, const std::vector<char>& blob, hipStream_t st) {
  psdp_handle_t h = s->h;
  if (blob.size() > a.bytes) {
    HIPCHK(hipStreamSynchronize(st));
    if (a.p) HIPCHK(hipFree(a.p));
    a.p = nullptr;
    a.bytes = 0;
    const size_t want = std::max<size_t>(blob.size() * 2, 1 << 20);
    if (hipMalloc(&a.p, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PSDP_ERR_NOMEM, "session descriptor area of %zu bytes", want);
    }
    a.bytes = want;
  }
  char* stage = nullptr;
  STCHK(staging_slot(h, blob.size(), &stage));
  memcpy(stage, blob.data(), blob.size());
  HIPCHK(hipMemcpyAsync(a.p, stage, blob.size(), hipMemcpyHostToDevice, st));
  STCHK(staging_release(h, st));
  return PSDP_OK;
}

// Enqueues the prologue of every block not yet started on the prologue
// stream (stream order keeps each upload behind the kernels of the previous
// prologue, so several may be queued).  Caller holds h->mu.
static psdp_status_t session_prologue(psdp_session_st* s) {
  psdp_handle_t h = s->h;
  std::vector<int> fresh;
  for (int i = 0; i < (int)s->blocks.size(); ++i)
    if (s->blocks[i].stage == -2) fresh.push_back(i);
  if (fresh.empty() || s->pros.size() >= 4) return PSDP_OK;
  std::vector<LanczosJob> lz;
  std::vector<PadJob> pj;
  std::vector<SplitJob> sj;
  std::vector<TileRef> padt, symt;
  std::vector<int2> rows, chunks;
  int max_m = 0;
  for (int g = 0; g < (int)fresh.size(); ++g) {
    const SessBlock& b = s->blocks[fresh[g]];
    lz.push_back(b.lz);
    pj.push_back(b.pj);
    sj.push_back(b.sj);
    const int T = b.q.npad / 64;
    for (int i = 0; i < T; ++i)
      for (int j = i; j < T; ++j) padt.push_back(TileRef{g, i, j, 0});
    const int nt = (b.q.npad + SYMV_TILE - 1) / SYMV_TILE;
    for (int j = 0; j < nt; ++j)
      for (int i = 0; i <= j; ++i) symt.push_back(TileRef{g, i, j, 0});
    for (int r = 0; r < nt; ++r) rows.push_back(make_int2(g, r));
    const int64_t nc = (int64_t)b.q.npad * b.q.npad / SPLIT_CHUNK;
    for (int64_t c = 0; c < nc; ++c) chunks.push_back(make_int2(g, (int)c));
    max_m = std::max(max_m, b.q.npad);
  }
  std::vector<char> blob;
  const size_t o_lz = blob_put(blob, lz.data(), lz.size());
  const size_t o_pj = blob_put(blob, pj.data(), pj.size());
  const size_t o_sj = blob_put(blob, sj.data(), sj.size());
  const size_t o_padt = blob_put(blob, padt.data(), padt.size());
  const size_t o_symt = blob_put(blob, symt.data(), symt.size());
  const size_t o_rows = blob_put(blob, rows.data(), rows.size());
  const size_t o_chunks = blob_put(blob, chunks.data(), chunks.size());
  // A region handed back by an earlier block is reused only after the host
  // saw that block's last tick complete, so no wait on the handle's stream.
  hipStream_t ps = s->pro_stream;
  STCHK(session_upload(s, s->pro_area, blob, ps));
  char* D = s->pro_area.p;
  HIPCHK(launch_sym_pad_batched((const PadJob*)(D + o_pj), (const TileRef*)(D + o_padt), (int)padt.size(), ps));
  HIPCHK(launch_lanczos_batched((const LanczosJob*)(D + o_lz), (int)fresh.size(), max_m, s->iters, s->tol, h->seed,
                                s->margin, s->tight, (const TileRef*)(D + o_symt), (int)symt.size(),
                                (const int2*)(D + o_rows), (int)rows.size(), ps));
  HIPCHK(launch_scale_split_batched((const SplitJob*)(D + o_sj), (const int2*)(D + o_chunks), (int)chunks.size(),
                                    s->sch.fmt[0], s->sch.final_fmt, ps));
  psdp_session_st::Prologue pr;
  if (!s->ev_pool.empty()) {
    pr.ev = s->ev_pool.back();
    s->ev_pool.pop_back();
  } else {
    HIPCHK(hipEventCreateWithFlags(&pr.ev, hipEventDisableTiming));
  }
  HIPCHK(hipEventRecord(pr.ev, ps));
  for (int i : fresh) {
    s->blocks[i].stage = -1;
    pr.tickets.push_back(s->blocks[i].ticket);
  }
  s->pros.push_back(std::move(pr));
  return PSDP_OK;
}

// Enqueues one tick over the ready blocks (stage >= 0).  Caller holds h->mu.
static psdp_status_t session_tick(psdp_session_st* s) {
  psdp_handle_t h = s->h;
  hipStream_t st = h->stream;
  struct Group {
    int fmt, avg, sub;
    std::vector<int> members;  // indices into s->blocks
    size_t o_desc = 0, o_tiles = 0;
    int ntiles = 0;
  };
  std::vector<Group> groups;
  for (int i = 0; i < (int)s->blocks.size(); ++i) {
    const SessBlock& b = s->blocks[i];
    if (b.stage < 0) continue;
    const int stg = b.stage;
    const bool fin = stg >= 3 * s->sch.nsteps;
    const int f = fin ? s->sch.final_fmt : s->sch.fmt[stg / 3];
    const int avg = fin ? s->sch.final_avg : (stg % 3 != 0 ? s->sch.avg[stg / 3] : 0);
    const int sub = block_tile(b.q.n, s->big_n) / 32;
    auto it = std::find_if(groups.begin(), groups.end(),
                           [&](const Group& g) { return g.fmt == f && g.avg == avg && g.sub == sub; });
    if (it == groups.end()) {
      groups.push_back(Group{f, avg, sub, {}});
      it = groups.end() - 1;
    }
    it->members.push_back(i);
  }
  if (groups.empty()) return PSDP_OK;
  std::vector<char> blob;
  for (auto& g : groups) {
    std::vector<ProdDesc> d;
    std::vector<int> npads, local;
    for (int k = 0; k < (int)g.members.size(); ++k) {
      const SessBlock& b = s->blocks[g.members[k]];
      d.push_back(b.pd[b.stage]);
      DESCCHK(d.back(), g.fmt);
      npads.push_back(b.q.npad);
      local.push_back(k);
    }
    const std::vector<TileRef> tl = batched_tile_order(local, 32 * g.sub, npads);
    g.o_desc = blob_put(blob, d.data(), d.size());
    g.o_tiles = blob_put(blob, tl.data(), tl.size());
    g.ntiles = (int)tl.size();
  }
  STCHK(session_upload(s, s->tick_area, blob, st));
  char* D = s->tick_area.p;
  for (const auto& g : groups) {
    if (!g.ntiles) continue;
    ProfScope pg(h, PSDP_PROF_PROD_F32 + g.fmt);
    const int grid = std::min(g.ntiles, g.sub >= 4 ? h->cus : 2 * h->cus);  // the resident workgroups
    HIPCHK(launch_sym_product_grouped(g.fmt, g.sub, g.avg, (const ProdDesc*)(D + g.o_desc),
                                      (const TileRef*)(D + g.o_tiles), g.ntiles, grid, st));
  }
  psdp_session_st::Tick t;
  if (!s->ev_pool.empty()) {
    t.ev = s->ev_pool.back();
    s->ev_pool.pop_back();
  } else {
    HIPCHK(hipEventCreateWithFlags(&t.ev, hipEventDisableTiming));
  }
  std::vector<SessBlock> keep;
  keep.reserve(s->blocks.size());
  for (auto& b : s->blocks) {
    if (b.stage >= 0 && ++b.stage >= s->nstages) t.finished.push_back(psdp_session_st::Done{b.ticket, b.off, b.bytes});
    else keep.push_back(std::move(b));
  }
  s->blocks.swap(keep);
  HIPCHK(hipEventRecord(t.ev, st));
  s->ticks.push_back(std::move(t));
  return PSDP_OK;
}

extern "C" {

psdp_status_t psdp_session_create(psdp_handle_t h, const psdp_config_t* cfg, size_t arena_bytes,
                                  psdp_session_t* out) {
  if (!h || !out) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  psdp_config_t def;
  if (!cfg) {
    psdp_default_config(&def, PSDP_METHOD_MIXED);
    cfg = &def;
  }
  if (!cfg->auto_scale) return fail(PSDP_ERR_UNSUPPORTED, "sessions project with the Lanczos bound (auto_scale = 1)");
  if (arena_bytes < (1 << 20)) return fail(PSDP_ERR_INVALID_ARG, "session arena below 1 MiB");
  auto* s = new psdp_session_st();
  s->h = h;
  const psdp_status_t sst = build_schedule(cfg, s->sch);
  if (sst != PSDP_OK) {
    delete s;
    return sst;
  }
  s->iters = cfg->lanczos_iters > 0 ? cfg->lanczos_iters : 20;
  if (s->iters > LANCZOS_MAX_ITER) {
    delete s;
    return fail(PSDP_ERR_INVALID_ARG, "at most %d Lanczos iterations are supported", LANCZOS_MAX_ITER);
  }
  s->tol = cfg->lanczos_tol > 0 ? cfg->lanczos_tol : 1e-10;
  s->margin = cfg->scale_margin > 0 ? cfg->scale_margin : 1.0;
  s->tight = cfg->bound_mode == PSDP_BOUND_TIGHT;
  s->nstages = 3 * s->sch.nsteps + 1;
  {
    std::lock_guard<std::mutex> lock(h->mu);
    s->big_n = h->batch_big_n > 0 ? h->batch_big_n : BIG_BLOCK_N;
    if (hipSetDevice(h->device) != hipSuccess || hipMalloc(&s->arena, arena_bytes) != hipSuccess) {
      (void)hipGetLastError();
      delete s;
      return fail(PSDP_ERR_NOMEM, "session arena of %zu bytes", arena_bytes);
    }
  }
  s->arena_bytes = arena_bytes;
  s->free_[0] = arena_bytes;
  if (hipStreamCreateWithFlags(&s->pro_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    if (s->pro_stream) (void)hipStreamDestroy(s->pro_stream);
    (void)hipFree(s->arena);
    delete s;
    return fail(PSDP_ERR_DEVICE, "session stream creation failed");
  }
  *out = s;
  return PSDP_OK;
}

psdp_status_t psdp_session_destroy(psdp_session_t s) {
  if (!s) return PSDP_OK;
  psdp_handle_t h = s->h;
  std::lock_guard<std::mutex> lock(h->mu);
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  if (s->pro_stream) (void)hipStreamSynchronize(s->pro_stream);
  for (auto& t : s->ticks) (void)hipEventDestroy(t.ev);
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  for (auto& p : s->pros) (void)hipEventDestroy(p.ev);
  if (s->pro_stream) (void)hipStreamDestroy(s->pro_stream);
  if (s->arena) (void)hipFree(s->arena);
  if (s->tick_area.p) (void)hipFree(s->tick_area.p);
  if (s->pro_area.p) (void)hipFree(s->pro_area.p);
  delete s;
  return PSDP_OK;
}

psdp_status_t psdp_session_block_bytes(psdp_session_t s, int64_t n, size_t* bytes) {
  if (!s || !bytes) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0 || n > 65536) return fail(PSDP_ERR_INVALID_ARG, "bad block size %lld", (long long)n);
  size_t so = 0;
  *bytes = session_block_bytes(s, (int)n, &so);
  return PSDP_OK;
}

psdp_status_t psdp_session_add(psdp_session_t s, const float* X, float* Y, int64_t n, int64_t* ticket) {
  if (!s || !X || !Y || !ticket) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0 || n > 65536) return fail(PSDP_ERR_INVALID_ARG, "bad block size %lld", (long long)n);
  psdp_handle_t h = s->h;
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  STCHK(session_retire(s));
  size_t so = 0;
  const size_t bytes = session_block_bytes(s, (int)n, &so);
  if (bytes > s->arena_bytes)
    return fail(PSDP_ERR_INVALID_ARG, "a block of n = %lld needs %zu bytes, more than the session arena (%zu)",
                (long long)n, bytes, s->arena_bytes);
  size_t off = 0;
  if (!session_alloc(s, bytes, &off))
    return fail(PSDP_ERR_NOMEM, "no room in the session arena for n = %lld now (%zu bytes)", (long long)n, bytes);
  SessBlock b;
  b.ticket = s->next_ticket++;
  b.off = off;
  b.bytes = bytes;
  layout_block(s->sch, (int)n, s->iters, s->big_n, off, b.q);
  b.st = (LanczosState*)(s->arena + off + so);
  b.pd.resize(s->nstages);
  block_descs(s->sch, b.q, s->arena, X, Y, b.st, b.pd.data(), 1, b.lz, b.pj, b.sj);
  b.stage = -2;
  *ticket = b.ticket;
  s->blocks.push_back(std::move(b));
  return PSDP_OK;
}

psdp_status_t psdp_session_step(psdp_session_t s, int* in_flight_ticks) {
  if (!s) return fail(PSDP_ERR_INVALID_ARG, "null session");
  psdp_handle_t h = s->h;
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  STCHK(session_retire(s));
  STCHK(session_prologue(s));
  STCHK(session_tick(s));
  if (in_flight_ticks) *in_flight_ticks = (int)s->ticks.size();
  return PSDP_OK;
}

psdp_status_t psdp_session_poll(psdp_session_t s, int64_t* tickets, int max, int* count, int* active,
                                int* in_flight_ticks) {
  if (!s || !count || (max > 0 && !tickets)) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  psdp_handle_t h = s->h;
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  STCHK(session_retire(s));
  const int k = std::min<int>(max, (int)s->completed.size());
  for (int i = 0; i < k; ++i) tickets[i] = s->completed[i];
  s->completed.erase(s->completed.begin(), s->completed.begin() + k);
  *count = k;
  if (active) {
    int a = (int)s->blocks.size();
    for (const auto& t : s->ticks) a += (int)t.finished.size();
    *active = a;
  }
  if (in_flight_ticks) *in_flight_ticks = (int)s->ticks.size();
  return PSDP_OK;
}

}  // extern "C"

psdp_status_t psdp_project_packed_f32(psdp_handle_t h, const float* X, float* Y, const int64_t* sizes, int count,
                                      const psdp_config_t* cfg) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (!X || !Y || !sizes) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (count <= 0 || count > 65535) return fail(PSDP_ERR_INVALID_ARG, "block count must be in [1, 65535]");
  std::vector<psdp_block_t> blocks((size_t)count);
  int64_t off = 0;
  for (int b = 0; b < count; ++b) {
    if (sizes[b] <= 0 || sizes[b] > 65536) return fail(PSDP_ERR_INVALID_ARG, "block %d: bad size", b);
    blocks[b] = psdp_block_t{X + off, Y + off, sizes[b]};
    off += sizes[b] * sizes[b];
  }
  return psdp_project_batched_f32(h, blocks.data(), count, cfg);
}

psdp_status_t psdp_autotune_batched(psdp_handle_t h, const psdp_block_t* blocks, int count,
                                    const psdp_config_t* cfg) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  psdp_config_t c;
  if (cfg) c = *cfg;
  else psdp_default_config(&c, PSDP_METHOD_MIXED);
  c.synchronize = 0;
  // candidate tile classes: the 128 x 128 class from n >= big (or none), the
  // two classes of a stage serial or on two streams
  const int cand[][2] = {{BIG_BLOCK_N, 0}, {1024, 0}, {768, 0}, {512, 0}, {512, 1}, {384, 1}, {256, 1}};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  int best = 0;
  float best_ms = 1e30f;
  for (int k = 0; k < (int)(sizeof cand / sizeof cand[0]); ++k) {
    STCHK(psdp_set_batched_tiling(h, cand[k][0], cand[k][1]));
    STCHK(psdp_project_batched_f32(h, blocks, count, &c));  // warm-up: workspace, tile lists
    constexpr int R = 2;
    HIPCHK(hipEventRecord(e0, h->stream));
    for (int r = 0; r < R; ++r) STCHK(psdp_project_batched_f32(h, blocks, count, &c));
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) best_ms = ms, best = k;
  }
  return psdp_set_batched_tiling(h, cand[best][0], cand[best][1]);
}

static psdp_status_t ensure_blas(psdp_handle_t h) {
  if (!h->blas) {
    h->blas = eig_create_blas();
    if (!h->blas) return fail(PSDP_ERR_DEVICE, "rocBLAS handle creation failed");
  }
  return PSDP_OK;
}

// Caller holds h->mu.
static psdp_status_t eig_locked(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues, int f64) {
  STCHK(ensure_blas(h));
  const size_t wsb = align256(eig_workspace_bytes((int)n, f64));
  void* base = nullptr;
  STCHK(arena_reserve(h, wsb + 256, &base));
  int* info = (int*)((char*)base + wsb);
  const int st = eig_project(h->blas, h->stream, mat, (int)n, f64, base, info, eigenvalues);
  if (st) return fail(PSDP_ERR_DEVICE, "rocSOLVER/rocBLAS call failed (status %d)", st);
  int info_h = 0;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&info_h, info, sizeof info_h, hipMemcpyDeviceToHost));
  if (info_h != 0) return fail(PSDP_ERR_NONFINITE, "syevd did not converge (info = %d)", info_h);
  return PSDP_OK;
}

static psdp_status_t eig_dropin(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues, int f64) {
  if (!h || !mat) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0 || n > 65536) return fail(PSDP_ERR_INVALID_ARG, "bad matrix size %lld", (long long)n);
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  return eig_locked(h, mat, n, eigenvalues, f64);
}

psdp_status_t psdp_eig_fp64_psd(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues) {
  return eig_dropin(h, mat, n, eigenvalues, 1);
}

psdp_status_t psdp_eig_fp32_psd(psdp_handle_t h, double* mat, int64_t n, double* eigenvalues) {
  return eig_dropin(h, mat, n, eigenvalues, 0);
}

// The reference returns a device array it allocated (src/eig_FP64_psd.cu:30,
// 46-49, 84); here it is allocated only when asked for, and never returned
// uninitialised.
static psdp_status_t eig_alloc(psdp_handle_t h, double* mat, int64_t n, double** out, int f64) {
  if (out) *out = nullptr;
  if (!h || !mat) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (n <= 0 || n > 65536) return fail(PSDP_ERR_INVALID_ARG, "bad matrix size %lld", (long long)n);
  double* w = nullptr;
  if (out) {
    HIPCHK(hipSetDevice(h->device));
    if (hipMalloc(&w, sizeof(double) * (size_t)n) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PSDP_ERR_NOMEM, "eigenvalue array of %lld doubles", (long long)n);
    }
  }
  const psdp_status_t st = eig_dropin(h, mat, n, w, f64);
  if (st != PSDP_OK) {
    if (w) (void)hipFree(w);
    return st;
  }
  if (out) *out = w;
  return PSDP_OK;
}

psdp_status_t psdp_eig_fp64_psd_alloc(psdp_handle_t h, double* mat, int64_t n, double** eigenvalues_out) {
  return eig_alloc(h, mat, n, eigenvalues_out, 1);
}

psdp_status_t psdp_eig_fp32_psd_alloc(psdp_handle_t h, double* mat, int64_t n, double** eigenvalues_out) {
  return eig_alloc(h, mat, n, eigenvalues_out, 0);
}

psdp_status_t psdp_free(void* p) {
  if (!p) return PSDP_OK;
  HIPCHK(hipFree(p));
  return PSDP_OK;
}

// MATLAB/psd_projection_MATLAB.cu:112-166: the five method strings of the
// MEX gateway (README.md:56), each an in-place float64 projection that
// returns when done; "mixed" and "mixed_mx" add the library's schedules.
psdp_status_t psdp_project_method(psdp_handle_t h, const char* method, double* mat, int64_t n,
                                  double* eigenvalues) {
  if (!method) return fail(PSDP_ERR_INVALID_ARG, "null method name");
  const std::string m(method);
  if (m == "composite_FP32") return psdp_composite_fp32_auto_scale(h, mat, n, nullptr, 0);
  if (m == "composite_FP16") return psdp_composite_fp16_auto_scale(h, mat, n, nullptr, 0);
  if (m == "composite_FP32_emulated") return psdp_composite_fp32_emulated_auto_scale(h, mat, n, nullptr, 0);
  if (m == "eig_FP64") return psdp_eig_fp64_psd(h, mat, n, eigenvalues);
  if (m == "eig_FP32") return psdp_eig_fp32_psd(h, mat, n, eigenvalues);
  if (m == "mixed") return dropin(h, mat, n, nullptr, 0, PSDP_METHOD_MIXED, 1);
  if (m == "mixed_mx") return dropin(h, mat, n, nullptr, 0, PSDP_METHOD_MIXED_MX, 1);
  return fail(PSDP_ERR_INVALID_ARG,
              "unknown method '%s' (composite_FP16, composite_FP32, composite_FP32_emulated, eig_FP64, eig_FP32, "
              "mixed, mixed_mx)",
              method);
}

// ---------------------------------------------------------------- eigenpairs and deflation
static psdp_status_t ensure_pinned(psdp_handle_t h) {
  if (!h->pinned && hipHostMalloc(&h->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    h->pinned = nullptr;
    return fail(PSDP_ERR_NOMEM, "pinned host allocation failed");
  }
  return PSDP_OK;
}

// Device memory that survives the projection inside a deflated call.
static psdp_status_t keep_reserve(psdp_handle_t h, size_t bytes, void** out) {
  if (bytes > h->keep_bytes) {
    if (h->keep) {
      HIPCHK(hipStreamSynchronize(h->stream));
      HIPCHK(hipFree(h->keep));
      h->keep = nullptr;
      h->keep_bytes = 0;
    }
    if (hipMalloc(&h->keep, bytes) != hipSuccess) {
      (void)hipGetLastError();
      h->keep = nullptr;
      return fail(PSDP_ERR_NOMEM, "eigenpair storage of %zu bytes failed", bytes);
    }
    h->keep_bytes = bytes;
  }
  *out = h->keep;
  return PSDP_OK;
}

static psdp_status_t eig_env(psdp_handle_t h, EigEnv* env) {
  STCHK(ensure_blas(h));
  STCHK(ensure_pinned(h));
  env->blas = h->blas;
  env->stream = h->stream;
  env->seed = h->seed;
  env->pinned = h->pinned;
  return PSDP_OK;
}

static psdp_status_t eigenpairs_locked(psdp_handle_t h, const double* A, int64_t n, int64_t k, int64_t* r,
                                       double* evals, double* evecs, bool upper_only, int64_t max_iter,
                                       double ortho_tol, bool normalize, double* bound) {
  if (max_iter == 0) max_iter = n;  // src/lanczos.cu:331-332
  if (upper_only) k = 0;
  EigEnv env;
  STCHK(eig_env(h, &env));
  void* base = nullptr;
  STCHK(arena_reserve(h, lanczos_pairs_ws_bytes(n, k, max_iter), &base));
  ProfScope ps(h, PSDP_PROF_EIGENPAIRS);
  int64_t rr = 0;
  STCHK(lanczos_pairs(env, A, n, k, max_iter, ortho_tol, h->ritz_row, normalize, upper_only, base, evals, evecs,
                      &rr, bound));
  if (r) *r = rr;
  return PSDP_OK;
}

static psdp_status_t lobpcg_locked(psdp_handle_t h, const double* A, double sign, double* V, double* D, int64_t n,
                                   int64_t m, bool warm, int maxiter, double tol) {
  EigEnv env;
  STCHK(eig_env(h, &env));
  void* base = nullptr;
  STCHK(arena_reserve(h, lobpcg_ws_bytes(n, m), &base));
  ProfScope ps(h, PSDP_PROF_LOBPCG);
  return lobpcg_run(env, A, sign, V, D, n, m, warm, maxiter, tol, base, &h->lobpcg_iters, &h->lobpcg_rnorm);
}

static psdp_status_t update_locked(psdp_handle_t h, double* X, int64_t n, const double* V, const double* lam,
                                   int64_t r, int64_t split, int mode) {
  if (r <= 0) return PSDP_OK;
  EigEnv env;
  STCHK(eig_env(h, &env));
  void* base = nullptr;
  STCHK(arena_reserve(h, align256(8 * (size_t)n * r), &base));
  ProfScope ps(h, PSDP_PROF_RANK_UPDATE);
  return rank_update(env, X, n, V, lam, (int)r, (int)split, mode, (double*)base);
}

// Eigenvectors (n x cols) followed by eigenvalues (cols) in the keep area.
static psdp_status_t keep_pairs(psdp_handle_t h, int64_t n, int64_t cols, double** V, double** lam) {
  cols = std::max<int64_t>(cols, 1);
  void* base = nullptr;
  const size_t vb = align256(8 * (size_t)n * cols);
  STCHK(keep_reserve(h, vb + align256(8 * (size_t)cols), &base));
  *V = (double*)base;
  *lam = (double*)((char*)base + vb);
  return PSDP_OK;
}

static bool size_ok(int64_t n) { return n > 0 && n <= 65536; }

psdp_status_t psdp_set_eigenpair_options(psdp_handle_t h, int residual_row, int normalize_vectors) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (residual_row != PSDP_RITZ_RESIDUAL_FIRST_ROW && residual_row != PSDP_RITZ_RESIDUAL_LAST_ROW)
    return fail(PSDP_ERR_INVALID_ARG, "unknown Ritz residual row %d", residual_row);
  std::lock_guard<std::mutex> lock(h->mu);
  h->ritz_row = residual_row;
  h->ritz_normalize = normalize_vectors != 0;
  return PSDP_OK;
}

psdp_status_t psdp_compute_eigenpairs(psdp_handle_t h, const double* A, int64_t n, int64_t k, int64_t* r,
                                      double* eigenvalues, double* eigenvectors, int upper_bound_only,
                                      int64_t max_iter, double tol, double ortho_tol, double* upper_bound) {
  (void)tol;  // the reference filters with ortho_tol only (src/lanczos.cu:437)
  if (!h || !A || !upper_bound) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (!size_ok(n)) return fail(PSDP_ERR_INVALID_ARG, "bad matrix size %lld", (long long)n);
  if (k < 0 || max_iter < 0) return fail(PSDP_ERR_INVALID_ARG, "negative k or max_iter");
  if (!upper_bound_only && (!r || (k > 0 && (!eigenvalues || !eigenvectors))))
    return fail(PSDP_ERR_INVALID_ARG, "null output array");
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  return eigenpairs_locked(h, A, n, k, upper_bound_only ? nullptr : r, eigenvalues, eigenvectors,
                           upper_bound_only != 0, max_iter, ortho_tol, h->ritz_normalize != 0, upper_bound);
}

psdp_status_t psdp_lobpcg(psdp_handle_t h, const double* A, double* V, double* D, int64_t n, int64_t m,
                          int warmstart, int maxiter, double tol) {
  if (!h || !A || !V || !D) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (!size_ok(n) || m <= 0 || 3 * m > n)
    return fail(PSDP_ERR_INVALID_ARG, "lobpcg needs 0 < 3 m <= n (n = %lld, m = %lld)", (long long)n,
                (long long)m);
  if (maxiter < 0) return fail(PSDP_ERR_INVALID_ARG, "negative maxiter");
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  return lobpcg_locked(h, A, 1.0, V, D, n, m, warmstart != 0, maxiter, tol);
}

psdp_status_t psdp_lobpcg_last(psdp_handle_t h, int* iterations, double* residual_norm) {
  if (!h) return fail(PSDP_ERR_INVALID_ARG, "null handle");
  if (iterations) *iterations = h->lobpcg_iters;
  if (residual_norm) *residual_norm = h->lobpcg_rnorm;
  return PSDP_OK;
}

psdp_status_t psdp_composite_fp32_auto_scale_deflate(psdp_handle_t h, double* mat, int64_t n, int64_t k,
                                                     double tol, double ortho_tol) {
  (void)tol;
  if (!h || !mat) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (!size_ok(n) || k < 0) return fail(PSDP_ERR_INVALID_ARG, "bad size (n = %lld, k = %lld)", (long long)n,
                                        (long long)k);
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  double *V = nullptr, *lam = nullptr;
  STCHK(keep_pairs(h, n, k, &V, &lam));
  int64_t r = 0;
  double bound = 0.0;
  // src/composite_FP32.cu:204-225: the k pairs of largest |lambda|, removed;
  // the kept Ritz vectors are normalised (the reference's are not: with a
  // basis that lost orthogonality its x x^T removes theta ||x||^2, which
  // leaves (||x||^2 - 1) theta behind in the result)
  if (k > 0) STCHK(eigenpairs_locked(h, mat, n, k, &r, lam, V, false, 0, ortho_tol, true, &bound));
  STCHK(update_locked(h, mat, n, V, lam, r, r, RANK_REMOVE));
  // :227-235: bound of the deflated matrix, s = 1.1 up + 1e-5
  double up = 0.0;
  STCHK(eigenpairs_locked(h, mat, n, 0, nullptr, nullptr, nullptr, true, 100, 1e-5, false, &up));
  const double scale = up > 0.0 ? 1.1 * up + 1e-5 : 1.0;
  // :237-241: s composite_FP32(X / s)
  psdp_config_t cfg;
  psdp_default_config(&cfg, PSDP_METHOD_COMPOSITE_FP32);
  cfg.auto_scale = 0;
  STCHK(project_locked(h, mat, 1, mat, 1, n, &cfg, nullptr, 0, scale));
  // :243-250: the positive ones added back
  STCHK(update_locked(h, mat, n, V, lam, r, r, RANK_ADD_POSITIVE));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PSDP_OK;
}

// LOBPCG deflation shared by the FP16 and FP64 deflated projections
// (src/composite_FP16.cu:300-333, src/eig_FP64_psd.cu:104-137): the k largest
// pairs are removed, then the k smallest of the remainder (as the largest of
// its negation).  V: n x 2k, lam: 2k (largest, then those of the negation).
static psdp_status_t lobpcg_deflate_locked(psdp_handle_t h, double* mat, int64_t n, int64_t k, int maxiter,
                                           double tol, double* V, double* lam) {
  if (k == 0) return PSDP_OK;
  STCHK(lobpcg_locked(h, mat, 1.0, V, lam, n, k, false, maxiter, tol));
  STCHK(update_locked(h, mat, n, V, lam, k, k, RANK_REMOVE));
  STCHK(lobpcg_locked(h, mat, -1.0, V + n * k, lam + k, n, k, false, maxiter, tol));
  STCHK(update_locked(h, mat, n, V + n * k, lam + k, k, 0, RANK_REMOVE));
  return PSDP_OK;
}

static psdp_status_t lobpcg_deflate_args(psdp_handle_t h, double* mat, int64_t n, int64_t k, int maxiter) {
  if (!h || !mat) return fail(PSDP_ERR_INVALID_ARG, "null argument");
  if (!size_ok(n) || k < 0 || 3 * k > n)
    return fail(PSDP_ERR_INVALID_ARG, "deflation by lobpcg needs 0 <= 3 k <= n (n = %lld, k = %lld)",
                (long long)n, (long long)k);
  if (maxiter < 0) return fail(PSDP_ERR_INVALID_ARG, "negative maxiter");
  return PSDP_OK;
}

psdp_status_t psdp_composite_fp16_auto_scale_deflate(psdp_handle_t h, double* mat, int64_t n, void* ws,
                                                     size_t ws_bytes, int64_t k, int maxiter, double tol) {
  STCHK(lobpcg_deflate_args(h, mat, n, k, maxiter));
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  double *V = nullptr, *lam = nullptr;
  STCHK(keep_pairs(h, n, 2 * k, &V, &lam));
  STCHK(lobpcg_deflate_locked(h, mat, n, k, maxiter, tol, V, lam));
  // src/composite_FP16.cu:335-346: Lanczos bound, s = 1.1 up + 1e-5
  double lo = 0.0, up = 0.0;
  {
    ProfScope ps(h, PSDP_PROF_LANCZOS);
    STCHK(two_norm_locked(h, mat, n, 20, 1e-10, &lo, &up));
  }
  const double scale = up > 0.0 ? 1.1 * up + 1e-5 : 1.0;
  psdp_config_t cfg;
  psdp_default_config(&cfg, PSDP_METHOD_COMPOSITE_FP16);
  cfg.auto_scale = 0;
  STCHK(project_locked(h, mat, 1, mat, 1, n, &cfg, ws, ws_bytes, scale));
  // :351-370
  STCHK(update_locked(h, mat, n, V, lam, 2 * k, k, RANK_ADD_POSITIVE));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PSDP_OK;
}

psdp_status_t psdp_eig_fp64_deflate(psdp_handle_t h, double* mat, int64_t n, int64_t k, int maxiter, double tol) {
  STCHK(lobpcg_deflate_args(h, mat, n, k, maxiter));
  std::lock_guard<std::mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  double *V = nullptr, *lam = nullptr;
  STCHK(keep_pairs(h, n, 2 * k, &V, &lam));
  STCHK(lobpcg_deflate_locked(h, mat, n, k, maxiter, tol, V, lam));
  STCHK(eig_locked(h, mat, n, nullptr, 1));  // src/eig_FP64_psd.cu:139-140
  STCHK(update_locked(h, mat, n, V, lam, 2 * k, k, RANK_ADD_POSITIVE));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PSDP_OK;
}

}  // extern "C"
