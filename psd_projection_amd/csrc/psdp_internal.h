// Types shared by the host driver and the HIP kernels of libpsdp.
#pragma once
#include <stdint.h>

namespace psdp {

// Operand formats of the symmetric products (values match psdp_format_t).
enum : int {
  FMT_F32 = 0,     // float32 operands on the FP32 MFMA
  FMT_F16 = 1,     // binary16 operands, float32 accumulation
  FMT_F16X3 = 2,   // hi + lo binary16 split of 2^8 v, 3 products, one accumulator
  FMT_BF16 = 3,    // bfloat16 operands
  FMT_BF16X3 = 4,  // hi + lo bfloat16 split, 3 products
  FMT_BF16X9 = 5,  // hi + mid + lo bfloat16 split, 9 products
  FMT_E4M3 = 6,    // 2^8 v rounded to the FP8 e4m3 grid, held in binary16 (binary16 MFMA)
  FMT_F16MX = 7,   // hi binary16 of 2^8 v + MX-scaled e4m3 copies of hi and lo for the cross terms
  FMT_COUNT = 8
};

// FMT_F16MX: w = 2^8 v, hi = f16(w), lo = w - hi (exact in float32).  The
// product is hi.hi on the binary16 MFMA plus the cross terms hi.lo + lo.hi
// from e4m3 copies on the scaled FP8 MFMA (v_mfma_scale_f32_32x32x64_f8f6f4),
// which runs at twice the binary16 rate: two binary16-MFMA equivalents per
// k instead of three.  Pieces (row stride ld, in 2-byte units):
//   [0] hi as binary16 at row * ld + col;
//   [1] e4m3 bytes: row r, 32-column block kb holds 64 bytes at byte
//       2 * r * ld + 64 * kb: the block's 32 hi values, then its 32 lo values;
//   [2] block scales (E8M0 exponents, value 2^(b - 127)), one uint16 per
//       (row, block) at kb * ld + r: low byte hi's, high byte lo's.
// A block's scale is 2^(E - 7) with E the exponent of its largest magnitude,
// so every scaled value is below 256 (no saturation); an all-zero block has
// scale 1.  Needs ld a multiple of 32.
constexpr int MX_BLOCK = 32;

// FMT_F16X3 splits s*v (s = 2^8, exact) into hi + lo binary16 pieces so that
// lo stays out of the binary16 subnormal range for the entry sizes of a
// matrix with spectrum in [-1, 1]; products carry s^2 and are rescaled by
// 2^-16 in the epilogue.  Operands must satisfy |v| < 2^8.
constexpr float F16X3_SCALE = 256.0f;
constexpr float F16X3_PROD_SCALE = 1.0f / 65536.0f;

// Epilogues of the symmetric product kernel.
enum : int {
  EPI_STORE = 0,       // C = X Y   -> c32 (optional) + operand pieces (out_fmt)
  EPI_POLY = 1,        // A = a A + b P3 + c X Y (in place) + pieces of A
  EPI_POLY_HALF_ID = 2,// as EPI_POLY, then S = (I + A)/2 written as pieces / c32
  EPI_OUTPUT = 3,      // out = scale * X Y on the logical n x n part (f64 or f32)
  EPI_PARTIAL = 4      // raw accumulators of one K chunk of a 256 x 256 tile -> part (split products)
};

// Split products (EPI_PARTIAL): K is cut into ksplit chunks of kchunk64 * 64
// elements (the last one shorter); tile list entry {slot, bi, bj, s} sums
// chunk s of tile slot from zero and stores the 256 x 256 raw accumulators,
// row-major, at part + (slot * ksplit + s) * 65536.  A second kernel adds
// the chunks of each tile in index order and runs the product's epilogue.

// One symmetric n_pad x n_pad product C = X Y of two symmetric matrices held
// as operand pieces (row-major view, leading dimension ld, zero padded).
struct ProdDesc {
  const void* x[3];  // pieces of X (uint16 or float)
  const void* y[3];  // pieces of Y
  int64_t ld;        // leading dimension of every padded buffer
  int npad;          // padded size (multiple of the tile)
  int n;             // logical size
  int mode;          // EPI_*
  int out_fmt;       // operand format written by EPI_STORE / EPI_POLY*
  float a, b, c;     // polynomial coefficients (EPI_POLY*)
  float* c32;        // float32 result (EPI_STORE) or A in/out (EPI_POLY*)
  const float* p3;   // float32 A^3 read by EPI_POLY* (binary16x3 products read A^3 from the pieces of x instead)
  uint16_t* outp[3]; // pieces written by EPI_STORE / EPI_POLY*
  void* out;         // EPI_OUTPUT destination (column-major n x n, ld = out_ld)
  int64_t out_ld;
  int out_f64;       // 1: double output, 0: float output
  const double* scale;  // EPI_OUTPUT multiplier (device scalar); null -> 1
  int* nonfinite;    // EPI_OUTPUT: set to 1 when a written value is not finite (null: unchecked)
  float* part;       // split products: partial tiles (EPI_PARTIAL writes, the reduction reads)
  int ksplit;        // split products: K chunks per tile
  int kchunk64;      // split products: chunk length in units of 64 k
};

// Tile of the upper triangle: matrix index + block coordinates (bi <= bj).
struct TileRef {
  int mat;
  int bi;
  int bj;
  int pad;
};

constexpr int LANCZOS_MAX_ITER = 64;

// Tile edge of the symmetric float32 GEMV of the Lanczos bound (also the rows
// per workgroup of its reduction), values per workgroup of the batched
// scale/split (a 64 x 64 tile holds a whole number of chunks).
constexpr int SYMV_TILE = 256;
constexpr int SPLIT_CHUNK = 4096;

// Device-resident Lanczos state (one per matrix).
struct LanczosState {
  double alpha[LANCZOS_MAX_ITER];
  double beta[LANCZOS_MAX_ITER];
  double u[LANCZOS_MAX_ITER];  // eigenvector of the largest Ritz value
  double beta_old;
  double theta;
  double lo, up;
  double scale;      // s = up > 0 ? up : 1 (times the caller's margin)
  double inv_scale;  // 1 / s
  double rnorm;
  int nb;            // completed iterations (nb_iter)
  int done;          // loop finished (break rule hit)
  int skip_ritz;     // nb == 0: no Ritz pair (all-zero input)
  int nonfinite;     // non-finite value met by the bound (input or bound)
  int out_nonfinite; // non-finite value written by the output product (diverged chain)
  int pad_;
};

// Device vectors of one Lanczos run (length m each; V is m x max_iter).
struct LanczosVecs {
  double* q;
  double* w;     // A q, then A A q
  double* w1;
  double* vold;
  double* V;     // column j at V + j*m
  double* y;
  double* ay;
  double* part;  // float32 GEMV partial sums: ceil(m / SYMV_TILE) x m (null: full-read GEMV only)
};

// One matrix of a batched Lanczos launch (float32 padded input).
struct LanczosJob {
  const float* A;
  int64_t ld;
  int n;         // logical size
  int m;         // padded size (vector length)
  int max_iter;  // min(iterations, n)
  int pad;
  LanczosState* st;
  LanczosVecs v;
};

// One matrix of a batched symmetrize/pad launch.
struct PadJob {
  const float* X;  // column-major n x n, leading dimension n
  int n;
  int npad;
  float* M;        // npad x npad
};

// One matrix of a batched scale/split launch.
struct SplitJob {
  float* a32;
  int64_t count;
  int64_t ld;  // row stride (FMT_F16MX blocks and scales)
  const double* inv;
  uint16_t* ap[3];
  float* x32;
  uint16_t* xp[3];
};

}  // namespace psdp
