// Input staging kernels: symmetrize + pad + convert, then scale + split into
// the operand pieces of the first filter step and of the final product.
//
// Replaces the reference's cublasDscal(1/up) on the FP64 matrix
// (src/composite_FP32.cu:182-184), convert_double_to_float
// (src/utils.cu:34-53, called at src/composite_FP32.cu:44,149) and the
// per-GEMM float4_to_half conversions of the operands
// (src/utils.cu:106-131, src/composite_FP16.cu:117,216-217).
#include "psdp_device.h"

namespace psdp {

// M = (X + X^T)/2 as float32, zero padded to npad x npad (row-major view,
// leading dimension ld >= npad; columns npad .. ld-1 are zeroed too, by the
// workgroups of the last column block, so that flat passes over npad * ld
// values see only finite data).  X is column-major n x n with leading
// dimension xld.
// One 64x64 tile pair (bi <= bj) per workgroup: both blocks are staged in
// LDS so that the transposed block is read and written coalesced.
// Columns npad .. ld-1 of the 64 rows of block row bi, written by the
// workgroup of the last column block (npad = 64 * gridDim.x).
__device__ __forceinline__ void zero_row_tail(float* __restrict__ M, int64_t ld, int bi, int bj) {
  const int64_t npad = 64 * (int64_t)gridDim.x;
  if (bj != (int)gridDim.x - 1 || ld == npad) return;
  const int extra = (int)(ld - npad);
  for (int e = threadIdx.x; e < 64 * extra; e += blockDim.x) {
    const int a = e / extra, b = e % extra;
    M[(int64_t)(bi * 64 + a) * ld + npad + b] = 0.0f;
  }
}

// s1[c][r] = X(i0 + r, j0 + c), s2[c][r] = X(j0 + r, i0 + c) for the 64 x 64
// tile pair (zero outside the n x n matrix; X column-major, leading dimension
// xld).  Every load reads a valid address (entry 0 stands in for those
// outside the matrix, n >= 1), so a thread issues all 32 of its loads before
// one wait; with the bound check on the load itself each one was a branch
// followed by its own wait.
template <typename TIn, typename TS>
__device__ __forceinline__ void load_tile_pair(const TIn* X, int64_t xld, int n, int i0, int j0,
                                               TS (*s1)[65], TS (*s2)[65]) {
  const int t = threadIdx.x, r = t & 63;
  const auto Xg = gbl(X);
  TIn v1[16], v2[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = (t >> 6) + 4 * k;
    const int gi = i0 + r, gj = j0 + c, hi = j0 + r, hj = i0 + c;
    v1[k] = Xg[(gi < n && gj < n) ? (int64_t)gi + (int64_t)gj * xld : 0];
    v2[k] = Xg[(hi < n && hj < n) ? (int64_t)hi + (int64_t)hj * xld : 0];
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int c = (t >> 6) + 4 * k;
    const int gi = i0 + r, gj = j0 + c, hi = j0 + r, hj = i0 + c;
    s1[c][r] = (gi < n && gj < n) ? (TS)v1[k] : (TS)0;
    s2[c][r] = (hi < n && hj < n) ? (TS)v2[k] : (TS)0;
  }
}

template <typename TIn>
__global__ __launch_bounds__(256) void sym_pad_kernel(const TIn* __restrict__ X, int64_t xld, int n,
                                                      float* __restrict__ M, int64_t ld) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  __shared__ float s1[64][65];
  __shared__ float s2[64][65];
  const int t = threadIdx.x;
  const int r = t & 63;
  const int i0 = bi * 64, j0 = bj * 64;
  load_tile_pair<TIn, float>(X, xld, n, i0, j0, s1, s2);
  __syncthreads();
  // M[i0 + a][j0 + b] = (X(i0+a, j0+b) + X(j0+b, i0+a)) / 2 = (s1[b][a] + s2[a][b]) / 2
#pragma unroll 4
  for (int a = t >> 6; a < 64; a += 4) {
    const int b = r;
    M[(int64_t)(i0 + a) * ld + j0 + b] = 0.5f * (s1[b][a] + s2[a][b]);
    if (bi != bj) M[(int64_t)(j0 + a) * ld + i0 + b] = 0.5f * (s2[b][a] + s1[a][b]);
  }
  zero_row_tail(M, ld, bi, bj);
}

// Symmetric inputs given in double are averaged in double before rounding.
template <>
__global__ __launch_bounds__(256) void sym_pad_kernel<double>(const double* __restrict__ X,
                                                              int64_t xld, int n,
                                                              float* __restrict__ M, int64_t ld) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  __shared__ double s1[64][65];
  __shared__ double s2[64][65];
  const int t = threadIdx.x;
  const int r = t & 63;
  const int i0 = bi * 64, j0 = bj * 64;
  load_tile_pair<double, double>(X, xld, n, i0, j0, s1, s2);
  __syncthreads();
#pragma unroll 4
  for (int a = t >> 6; a < 64; a += 4) {
    const int b = r;
    M[(int64_t)(i0 + a) * ld + j0 + b] = (float)(0.5 * (s1[b][a] + s2[a][b]));
    if (bi != bj) M[(int64_t)(j0 + a) * ld + i0 + b] = (float)(0.5 * (s2[b][a] + s1[a][b]));
  }
  zero_row_tail(M, ld, bi, bj);
}

struct SplitArgs {
  float* a32;            // in: unscaled M; out: scaled iterate A0 (float32)
  int64_t count;         // npad * ld (a multiple of 4096)
  int64_t ld;            // row stride (FMT_F16MX blocks and scales)
  const double* inv;     // device scalar 1/s (null -> 1)
  int fmt_a;             // operand format of filter step 0
  uint16_t* ap[3];
  int fmt_x;             // operand format of the final product's X operand
  float* x32;            // FMT_F32 copy of X_s
  uint16_t* xp[3];
};

// Four consecutive values of one row (index i, a multiple of 4) in format
// fmt; FMT_F16MX needs the eight lanes of each 32-value block together.
__device__ __forceinline__ void store_operand4(int fmt, const float v[4], int64_t i, float* f32,
                                               uint16_t* const* pieces);
__device__ __forceinline__ void store_quad(int fmt, const float v[4], int64_t i, int64_t ld, float* f32,
                                           uint16_t* const* pieces) {
  if (fmt == FMT_F16MX) {
    uint16_t hi[4];
    uint32_t h8, l8, sc;
    mx_split4(v, hi, h8, l8, sc);
    uint16_t* const pc[3] = {pieces[0], pieces[1], pieces[2]};
    mx_store4(pc, i / ld, (int)(i % ld), ld, hi, h8, l8, sc);
    return;
  }
  store_operand4(fmt, v, i, f32, pieces);
}

// One quad per thread and step; a wave's 64 quads are in range together
// (count is a multiple of 4096), so the FMT_F16MX lane groups stay whole.
__global__ __launch_bounds__(256) void scale_split_kernel(SplitArgs s) {
  const double inv = s.inv ? *s.inv : 1.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t quads = s.count / 4;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < quads; t += stride) {
    const int64_t i = 4 * t;
    const f32x4 a = *(const __attribute__((address_space(1))) f32x4*)gbl(s.a32 + i);
    const float v[4] = {(float)((double)a[0] * inv), (float)((double)a[1] * inv), (float)((double)a[2] * inv),
                        (float)((double)a[3] * inv)};
    *(__attribute__((address_space(1))) f32x4*)gbl(s.a32 + i) = f32x4{v[0], v[1], v[2], v[3]};
    if (s.fmt_a != FMT_F32) store_quad(s.fmt_a, v, i, s.ld, nullptr, s.ap);
    store_quad(s.fmt_x, v, i, s.ld, s.x32, s.xp);
  }
}

// Batched form (float32 inputs): workgroup g takes the 64 x 64 tile pair
// tiles[g] (matrix, bi <= bj) from a flat list over every matrix.
__global__ __launch_bounds__(256) void sym_pad_batched_kernel(const PadJob* __restrict__ jobs,
                                                              const TileRef* __restrict__ tiles) {
  const TileRef tr = tiles[blockIdx.x];
  const PadJob& J = jobs[tr.mat];
  const int bi = tr.bi, bj = tr.bj;
  __shared__ float s1[64][65];
  __shared__ float s2[64][65];
  const int t = threadIdx.x;
  const int r = t & 63;
  const int i0 = bi * 64, j0 = bj * 64;
  const int n = J.n;
  const int64_t ld = J.npad;
  const auto Mg = gbl(J.M);
  load_tile_pair<float, float>(J.X, n, n, i0, j0, s1, s2);
  __syncthreads();
#pragma unroll 4
  for (int a = t >> 6; a < 64; a += 4) {
    const int b = r;
    Mg[(int64_t)(i0 + a) * ld + j0 + b] = 0.5f * (s1[b][a] + s2[a][b]);
    if (bi != bj) Mg[(int64_t)(j0 + a) * ld + i0 + b] = 0.5f * (s2[b][a] + s1[a][b]);
  }
}

// Stores four consecutive values (index i, a multiple of 4) in format fmt
// with one vector store per piece.
__device__ __forceinline__ void store_operand4(int fmt, const float v[4], int64_t i, float* f32,
                                               uint16_t* const* pieces) {
  if (fmt == FMT_F32) {
    *(__attribute__((address_space(1))) f32x4*)gbl(f32 + i) = f32x4{v[0], v[1], v[2], v[3]};
    return;
  }
  uint16_t p[4][3];
#pragma unroll
  for (int e = 0; e < 4; ++e) split_value(fmt, v[e], p[e]);
  const int np = fmt_pieces(fmt);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q >= np) break;
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 w = {(uint32_t)p[0][q] | ((uint32_t)p[1][q] << 16), (uint32_t)p[2][q] | ((uint32_t)p[3][q] << 16)};
    *(__attribute__((address_space(1))) u32x2*)gbl(pieces[q] + i) = w;
  }
}

// Workgroup g scales and splits the SPLIT_CHUNK values chunks[g].y of matrix
// chunks[g].x (every matrix holds a multiple of 64 x 64 values).
__global__ __launch_bounds__(256) void scale_split_batched_kernel(const SplitJob* __restrict__ jobs,
                                                                  const int2* __restrict__ chunks, int fmt_a,
                                                                  int fmt_x) {
  const int2 ck = chunks[blockIdx.x];
  const SplitJob& J = jobs[ck.x];
  const double inv = J.inv ? *J.inv : 1.0;
  uint16_t* ap[3] = {J.ap[0], J.ap[1], J.ap[2]};
  uint16_t* xp[3] = {J.xp[0], J.xp[1], J.xp[2]};
  const int64_t base = (int64_t)ck.y * SPLIT_CHUNK;
#pragma unroll
  for (int r = 0; r < SPLIT_CHUNK / 1024; ++r) {
    const int64_t i = base + (int64_t)r * 1024 + 4 * threadIdx.x;
    const f32x4 a = *(const __attribute__((address_space(1))) f32x4*)gbl(J.a32 + i);
    const float v[4] = {(float)((double)a[0] * inv), (float)((double)a[1] * inv), (float)((double)a[2] * inv),
                        (float)((double)a[3] * inv)};
    *(__attribute__((address_space(1))) f32x4*)gbl(J.a32 + i) = f32x4{v[0], v[1], v[2], v[3]};
    if (fmt_a != FMT_F32) store_quad(fmt_a, v, i, J.ld, nullptr, ap);
    store_quad(fmt_x, v, i, J.ld, J.x32, xp);
  }
}

// Synthetic symmetric matrix for the autotuner: M(i, j) = (2 u - 1)/sqrt(n)
// with u from the seeded stream at index min(i,j) * n + max(i,j) for
// i, j < n, zero elsewhere (columns up to ld).
__global__ __launch_bounds__(256) void fill_sym_kernel(float* __restrict__ M, int64_t ld, int n, int npad,
                                                       uint64_t seed) {
  const int64_t total = (int64_t)npad * ld;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float sc = 1.0f / sqrtf((float)n);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e / ld), j = (int)(e % ld);
    float v = 0.0f;
    if (i < n && j < n) {
      const int a = i < j ? i : j, b = i < j ? j : i;
      v = (2.0f * (float)uniform01(seed, (uint64_t)a * n + b) - 1.0f) * sc;
    }
    M[e] = v;
  }
}

// dst = src for `words` 8-byte words (device to device, as a kernel so that a
// captured graph holds no copy node).
__global__ void copy_words_kernel(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_words(void* dst, const void* src, int64_t words, hipStream_t st) {
  int64_t blocks = (words + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(copy_words_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint64_t*)dst,
                     (const uint64_t*)src, words);
  return hipGetLastError();
}

hipError_t launch_fill_sym(float* M, int64_t ld, int n, int npad, uint64_t seed, hipStream_t st) {
  int64_t blocks = ((int64_t)npad * ld + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_sym_kernel, dim3((unsigned)blocks), dim3(256), 0, st, M, ld, n, npad, seed);
  return hipGetLastError();
}

hipError_t launch_sym_pad_batched(const PadJob* jobs, const TileRef* tiles, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(sym_pad_batched_kernel, dim3(ntiles), dim3(256), 0, st, jobs, tiles);
  return hipGetLastError();
}

hipError_t launch_scale_split_batched(const SplitJob* jobs, const int2* chunks, int nchunks, int fmt_a, int fmt_x,
                                      hipStream_t st) {
  hipLaunchKernelGGL(scale_split_batched_kernel, dim3(nchunks), dim3(256), 0, st, jobs, chunks, fmt_a, fmt_x);
  return hipGetLastError();
}

hipError_t launch_sym_pad(const void* X, int x_f64, int64_t xld, int n, float* M, int64_t ld, int npad,
                          hipStream_t st) {
  const int T = npad / 64;
  dim3 grid(T, T);
  if (x_f64)
    hipLaunchKernelGGL(sym_pad_kernel<double>, grid, dim3(256), 0, st, (const double*)X, xld, n, M, ld);
  else
    hipLaunchKernelGGL(sym_pad_kernel<float>, grid, dim3(256), 0, st, (const float*)X, xld, n, M, ld);
  return hipGetLastError();
}

hipError_t launch_scale_split(float* a32, int64_t count, int64_t ld, const double* inv, int fmt_a,
                              uint16_t* const ap[3], int fmt_x, float* x32, uint16_t* const xp[3], hipStream_t st) {
  if (count % 4096 != 0 || ld % 32 != 0) return hipErrorInvalidValue;
  SplitArgs s;
  s.a32 = a32;
  s.count = count;
  s.ld = ld;
  s.inv = inv;
  s.fmt_a = fmt_a;
  s.fmt_x = fmt_x;
  s.x32 = x32;
  for (int i = 0; i < 3; ++i) {
    s.ap[i] = ap[i];
    s.xp[i] = xp[i];
  }
  int64_t blocks = (count / 4 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(scale_split_kernel, dim3((unsigned)blocks), dim3(256), 0, st, s);
  return hipGetLastError();
}

}  // namespace psdp
