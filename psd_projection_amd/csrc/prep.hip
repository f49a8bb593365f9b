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
  // s1[c][r] = X(i0 + r, j0 + c); s2[c][r] = X(j0 + r, i0 + c)
#pragma unroll 4
  for (int c = t >> 6; c < 64; c += 4) {
    const int gi = i0 + r, gj = j0 + c;
    s1[c][r] = (gi < n && gj < n) ? (float)X[(int64_t)gi + (int64_t)gj * xld] : 0.0f;
    const int hi = j0 + r, hj = i0 + c;
    s2[c][r] = (hi < n && hj < n) ? (float)X[(int64_t)hi + (int64_t)hj * xld] : 0.0f;
  }
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
#pragma unroll 4
  for (int c = t >> 6; c < 64; c += 4) {
    const int gi = i0 + r, gj = j0 + c;
    s1[c][r] = (gi < n && gj < n) ? X[(int64_t)gi + (int64_t)gj * xld] : 0.0;
    const int hi = j0 + r, hj = i0 + c;
    s2[c][r] = (hi < n && hj < n) ? X[(int64_t)hi + (int64_t)hj * xld] : 0.0;
  }
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
  int64_t count;         // npad * npad
  const double* inv;     // device scalar 1/s (null -> 1)
  int fmt_a;             // operand format of filter step 0
  uint16_t* ap[3];
  int fmt_x;             // operand format of the final product's X operand
  float* x32;            // FMT_F32 copy of X_s
  uint16_t* xp[3];
};

__global__ __launch_bounds__(256) void scale_split_kernel(SplitArgs s) {
  const double inv = s.inv ? *s.inv : 1.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.count; i += stride) {
    const float v = (float)((double)s.a32[i] * inv);
    s.a32[i] = v;
    if (s.fmt_a != FMT_F32) store_operand(s.fmt_a, v, i, nullptr, s.ap);
    store_operand(s.fmt_x, v, i, s.x32, s.xp);
  }
}

// Batched form (float32 inputs): blockIdx.z selects the matrix; tiles beyond
// that matrix's padded size exit at once.
__global__ __launch_bounds__(256) void sym_pad_batched_kernel(const PadJob* __restrict__ jobs) {
  const PadJob& J = jobs[blockIdx.z];
  const int T = J.npad / 64;
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi >= T || bj >= T || bi > bj) return;
  __shared__ float s1[64][65];
  __shared__ float s2[64][65];
  const int t = threadIdx.x;
  const int r = t & 63;
  const int i0 = bi * 64, j0 = bj * 64;
  const int n = J.n;
  const int64_t ld = J.npad;
#pragma unroll 4
  for (int c = t >> 6; c < 64; c += 4) {
    const int gi = i0 + r, gj = j0 + c;
    s1[c][r] = (gi < n && gj < n) ? J.X[(int64_t)gi + (int64_t)gj * n] : 0.0f;
    const int hi = j0 + r, hj = i0 + c;
    s2[c][r] = (hi < n && hj < n) ? J.X[(int64_t)hi + (int64_t)hj * n] : 0.0f;
  }
  __syncthreads();
#pragma unroll 4
  for (int a = t >> 6; a < 64; a += 4) {
    const int b = r;
    J.M[(int64_t)(i0 + a) * ld + j0 + b] = 0.5f * (s1[b][a] + s2[a][b]);
    if (bi != bj) J.M[(int64_t)(j0 + a) * ld + i0 + b] = 0.5f * (s2[b][a] + s1[a][b]);
  }
}

__global__ __launch_bounds__(256) void scale_split_batched_kernel(const SplitJob* __restrict__ jobs, int fmt_a,
                                                                  int fmt_x) {
  const SplitJob& J = jobs[blockIdx.y];
  const double inv = J.inv ? *J.inv : 1.0;
  uint16_t* ap[3] = {J.ap[0], J.ap[1], J.ap[2]};
  uint16_t* xp[3] = {J.xp[0], J.xp[1], J.xp[2]};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < J.count; i += stride) {
    const float v = (float)((double)J.a32[i] * inv);
    J.a32[i] = v;
    if (fmt_a != FMT_F32) store_operand(fmt_a, v, i, nullptr, ap);
    store_operand(fmt_x, v, i, J.x32, xp);
  }
}

hipError_t launch_sym_pad_batched(const PadJob* jobs, int count, int max_npad, hipStream_t st) {
  const int T = max_npad / 64;
  hipLaunchKernelGGL(sym_pad_batched_kernel, dim3(T, T, count), dim3(256), 0, st, jobs);
  return hipGetLastError();
}

hipError_t launch_scale_split_batched(const SplitJob* jobs, int count, int64_t max_count, int fmt_a, int fmt_x,
                                      hipStream_t st) {
  int64_t blocks = (max_count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(scale_split_batched_kernel, dim3((unsigned)blocks, count), dim3(256), 0, st, jobs, fmt_a,
                     fmt_x);
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

hipError_t launch_scale_split(float* a32, int64_t count, const double* inv, int fmt_a, uint16_t* const ap[3],
                              int fmt_x, float* x32, uint16_t* const xp[3], hipStream_t st) {
  SplitArgs s;
  s.a32 = a32;
  s.count = count;
  s.inv = inv;
  s.fmt_a = fmt_a;
  s.fmt_x = fmt_x;
  s.x32 = x32;
  for (int i = 0; i < 3; ++i) {
    s.ap[i] = ap[i];
    s.xp[i] = xp[i];
  }
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(scale_split_kernel, dim3((unsigned)blocks), dim3(256), 0, st, s);
  return hipGetLastError();
}

}  // namespace psdp
