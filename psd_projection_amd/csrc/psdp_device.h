// Device-side definitions shared by the HIP kernels of libpsdp (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "psdp_internal.h"

namespace psdp {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Global-memory view of a pointer that was read from device memory (a field
// of a launch descriptor).  Without it the compiler cannot prove the address
// space and emits FLAT accesses, which also count on the LDS counter: every
// wait for an LDS read would then wait for all outstanding stores as well.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// ---------------------------------------------------------------------------
// Scalar conversions (round to nearest even).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) {
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_bits_to_f32(uint16_t b) {
  return (float)__builtin_bit_cast(_Float16, b);
}
// Value of a binary16x3 operand entry from its pieces, (hi + lo) 2^-8: within
// 2^-22 |v| + 2^-33 of the float32 value v it was split from (the absolute
// term where lo falls into the binary16 subnormals).
__device__ __forceinline__ float f16x3_value(uint16_t hi, uint16_t lo) {
  return (f16_bits_to_f32(hi) + f16_bits_to_f32(lo)) * (1.0f / F16X3_SCALE);
}
// Branch-free: NaN keeps its sign and top payload bits with the quiet bit set;
// infinities and finite values round to nearest even.
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  const uint32_t q = (u >> 16) | 0x40u;
  return (uint16_t)((u & 0x7fffffffu) > 0x7f800000u ? q : r);
}
__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) {
  return __uint_as_float(((uint32_t)b) << 16);
}

// Nearest value of the FP8 e4m3 grid (OCP E4M3FN: 3 mantissa bits, normal
// exponents down to 2^-6, subnormal quantum 2^-9, largest finite 448; values
// beyond saturate), ties to even.  Exact in binary16.
__device__ __forceinline__ float round_e4m3(float w) {
  const float a = fabsf(w);
  if (!(a < 448.0f)) return a == a ? copysignf(448.0f, w) : w;  // saturate (NaN stays NaN)
  int e;
  frexpf(a, &e);                      // a = m 2^e, m in [0.5, 1)
  const int q = (e - 1 < -6 ? -6 : e - 1) - 3;
  return ldexpf(rintf(ldexpf(w, -q)), q);
}

// Splits a float32 value into the pieces of operand format `fmt`.
// F16X3: w = 2^8 v (exact), hi = f16(w), lo = f16(w - hi).  BF16X3: hi, v - hi.
// BF16X9: hi, mid, lo (three bfloat16 terms).
__device__ __forceinline__ void split_value(int fmt, float v, uint16_t out[3]) {
  switch (fmt) {
    case FMT_F16: out[0] = f32_to_f16_bits(v); break;
    case FMT_F16X3: {
      const float w = v * F16X3_SCALE;
      uint16_t h = f32_to_f16_bits(w);
      out[0] = h;
      out[1] = f32_to_f16_bits(w - f16_bits_to_f32(h));
      break;
    }
    case FMT_BF16: out[0] = f32_to_bf16_bits(v); break;
    case FMT_BF16X3: {
      uint16_t h = f32_to_bf16_bits(v);
      out[0] = h;
      out[1] = f32_to_bf16_bits(v - bf16_bits_to_f32(h));
      break;
    }
    case FMT_E4M3: out[0] = f32_to_f16_bits(round_e4m3(v * F16X3_SCALE)); break;
    case FMT_BF16X9: {
      uint16_t h = f32_to_bf16_bits(v);
      float r = v - bf16_bits_to_f32(h);
      uint16_t m = f32_to_bf16_bits(r);
      out[0] = h;
      out[1] = m;
      out[2] = f32_to_bf16_bits(r - bf16_bits_to_f32(m));
      break;
    }
    default: break;
  }
}

// ---------------------------------------------------------------------------
// FMT_F16MX pieces (layout in psdp_internal.h).  Eight consecutive lanes
// hold the four-value quads of one 32-value block; every lane of the eight
// must call these together (the block maxima go through lane shuffles).
// ---------------------------------------------------------------------------
// E8M0 scale of a block whose largest magnitude has float bits m (magnitudes
// compare as unsigned integers): 2^(E - 7) for exponent field E, so the
// block's values scale below 256; an all-zero block gets 2^0.
__device__ __forceinline__ uint32_t mx_scale_byte(uint32_t m) {
  if (m == 0) return 127u;
  const int b = (int)(m >> 23) - 7;
  return (uint32_t)(b < 0 ? 0 : b);
}
// 2^(127 - b): divides by the scale exactly (b <= 248 for finite blocks).
__device__ __forceinline__ float mx_inv_scale(uint32_t b) { return __uint_as_float((254u - b) << 23); }

__device__ __forceinline__ uint32_t max_of_8_lanes(uint32_t v) {
  v = max(v, (uint32_t)__shfl_xor((int)v, 1, 64));
  v = max(v, (uint32_t)__shfl_xor((int)v, 2, 64));
  v = max(v, (uint32_t)__shfl_xor((int)v, 4, 64));
  return v;
}

// Four consecutive values v (of one row, column a multiple of 4) to their
// binary16 hi pieces, their packed e4m3 bytes of hi and lo (h8, l8: value e
// in byte e) and the block's scale pair (hi's | lo's << 8).
__device__ __forceinline__ void mx_split4(const float v[4], uint16_t hi[4], uint32_t& h8, uint32_t& l8,
                                          uint32_t& sc) {
  float w[4], lo[4];
  uint32_t mw = 0, ml = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    w[e] = v[e] * F16X3_SCALE;
    hi[e] = f32_to_f16_bits(w[e]);
    lo[e] = w[e] - f16_bits_to_f32(hi[e]);
    mw = max(mw, __float_as_uint(w[e]) & 0x7fffffffu);
    ml = max(ml, __float_as_uint(lo[e]) & 0x7fffffffu);
  }
  const uint32_t bw = mx_scale_byte(max_of_8_lanes(mw)), bl = mx_scale_byte(max_of_8_lanes(ml));
  const float iw = mx_inv_scale(bw), il = mx_inv_scale(bl);
  int ph = 0, pl = 0;
  ph = __builtin_amdgcn_cvt_pk_fp8_f32(w[0] * iw, w[1] * iw, ph, false);
  ph = __builtin_amdgcn_cvt_pk_fp8_f32(w[2] * iw, w[3] * iw, ph, true);
  pl = __builtin_amdgcn_cvt_pk_fp8_f32(lo[0] * il, lo[1] * il, pl, false);
  pl = __builtin_amdgcn_cvt_pk_fp8_f32(lo[2] * il, lo[3] * il, pl, true);
  h8 = (uint32_t)ph;
  l8 = (uint32_t)pl;
  sc = bw | (bl << 8);
}

// Byte offset of element (r, c) in piece 1 of FMT_F16MX (hi's byte; lo's is
// 32 bytes further), and the index of its block's scale in piece 2.
__device__ __forceinline__ int64_t mx_byte(int64_t r, int c, int64_t ld) {
  return 2 * r * ld + 64 * (c >> 5) + (c & 31);
}
__device__ __forceinline__ int64_t mx_scale_index(int64_t r, int c, int64_t ld) { return (int64_t)(c >> 5) * ld + r; }

// Stores the quad of mx_split4 at (r, c .. c + 3); the lane holding the
// block's first quad also stores the scale pair.
__device__ __forceinline__ void mx_store4(uint16_t* const pieces[3], int64_t r, int c, int64_t ld,
                                          const uint16_t hi[4], uint32_t h8, uint32_t l8, uint32_t sc) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 pk = {(uint32_t)hi[0] | ((uint32_t)hi[1] << 16), (uint32_t)hi[2] | ((uint32_t)hi[3] << 16)};
  *(__attribute__((address_space(1))) u32x2*)gbl(pieces[0] + r * ld + c) = pk;
  char* b8 = (char*)pieces[1] + mx_byte(r, c, ld);
  *(__attribute__((address_space(1))) uint32_t*)gbl((uint32_t*)b8) = h8;
  *(__attribute__((address_space(1))) uint32_t*)gbl((uint32_t*)(b8 + 32)) = l8;
  if ((c & 31) == 0) *gbl(pieces[2] + mx_scale_index(r, c, ld)) = (uint16_t)sc;
}

__device__ __forceinline__ int fmt_pieces(int fmt) {
  return fmt == FMT_F16X3 || fmt == FMT_BF16X3 ? 2 : (fmt == FMT_BF16X9 ? 3 : 1);
}

// Stores value v of element idx in operand format fmt.  For FMT_F32 the
// value goes to f32 (piece pointers unused).
__device__ __forceinline__ void store_operand(int fmt, float v, int64_t idx, float* f32,
                                              uint16_t* const* pieces) {
  if (fmt == FMT_F32) {
    f32[idx] = v;
    return;
  }
  uint16_t p[3];
  split_value(fmt, v, p);
  const int np = fmt_pieces(fmt);
  pieces[0][idx] = p[0];
  if (np > 1) pieces[1][idx] = p[1];
  if (np > 2) pieces[2][idx] = p[2];
}


// ---------------------------------------------------------------------------
// Seeded counter-based uniform stream: element i of stream `seed` is
// (splitmix64(seed ^ C (i + 1)) >> 11 + 1) 2^-53, a value in (0, 1].
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t i) {
  return ((double)(splitmix64(seed ^ (0x632be59bd9b4e019ull * (i + 1))) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Wave / block reductions (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the whole block; every thread receives the result.  `red` must
// hold blockDim.x/64 doubles of shared memory.  Deterministic order.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

}  // namespace psdp
