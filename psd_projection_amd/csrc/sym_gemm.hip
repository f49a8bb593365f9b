// Symmetric-output MFMA products C = X Y of two symmetric matrices on gfx950,
// with the composite-filter update fused into the epilogue.
//
// Replaces the cuBLAS calls of one filter step of the reference
// (src/composite_FP32.cu:103-118: Sgemm A2 = A A, Sgemm A3 = A2 A, Sscal,
// Saxpy, Sgemm A += c A3 A2, symmetrizeFloat; src/composite_FP16.cu:117-166
// for FP16 operands; src/composite_FP32_emulated.cu:108-155 for the
// emulated-FP32 products) and the final W = X S product (:143-156).
//
// Layout: every operand is a zero-padded n_pad x n_pad symmetric matrix in a
// row-major view with leading dimension ld.  Because X and Y are symmetric,
// C[i][j] = sum_k X[i][k] Y[j][k]: both operands are read along contiguous k,
// 16 bytes per lane.  Only tiles with bi <= bj are computed; each result is
// written at (i, j) and (j, i), so every output is exactly symmetric (this
// replaces the reference's explicit symmetrize passes).  The averaged form
// (AVG) accumulates C and C^T on the same tile and stores (C + C^T)/2.
//
// Tiling: 256 threads = 2 x 2 waves; each wave owns a (16*SUB) x (16*SUB)
// sub-block made of 16x16 MFMA tiles (SUB = 8 gives 256 x 256 block tiles
// with 128 x 128 per wave, the shape used for large n; SUB = 4 and 2 give
// 128 and 64).  The two-piece formats run their 256 x 256 tiles on 8 waves of
// 128 x 64 each instead (sym_tile_dma8).  K is staged through LDS in BK-wide slabs, double buffered
// (one barrier per slab).  Rows are padded by 16 bytes so that the 16-byte
// fragment reads of 16 consecutive rows hit distinct banks.  Split formats
// (x3, x9) accumulate every partial product into one accumulator.
#include "psdp_device.h"

namespace psdp {

template <int FMT>
struct FmtTraits;
template <>
struct FmtTraits<FMT_F32> {
  using E = float;
  static constexpr int P = 1, BK = 32;
};
template <>
struct FmtTraits<FMT_F16> {
  using E = uint16_t;
  static constexpr int P = 1, BK = 64;
};
template <>
struct FmtTraits<FMT_F16X3> {
  using E = uint16_t;
  static constexpr int P = 2, BK = 32;
};
template <>
struct FmtTraits<FMT_BF16> {
  using E = uint16_t;
  static constexpr int P = 1, BK = 64;
};
template <>
struct FmtTraits<FMT_BF16X3> {
  using E = uint16_t;
  static constexpr int P = 2, BK = 32;
};
template <>
struct FmtTraits<FMT_E4M3> {
  using E = uint16_t;
  static constexpr int P = 1, BK = 64;
};
template <>
struct FmtTraits<FMT_BF16X9> {
  using E = uint16_t;
  static constexpr int P = 3, BK = 32;
};
// Two 2-byte-per-k pieces per panel (hi binary16, the e4m3 hi/lo bytes) staged
// like FP16x3's, plus the block scales (piece 2) of every slab.
template <>
struct FmtTraits<FMT_F16MX> {
  using E = uint16_t;
  static constexpr int P = 2, BK = 32;
};

template <int FMT, int SUB, bool AVG = false>
struct Tiling {
  using F = FmtTraits<FMT>;
  using E = typename F::E;
  static constexpr int NPROD = AVG ? 2 : 1;    // averaged products keep C and C^T
  static constexpr int ESZ = (int)sizeof(E);
  static constexpr int WT = 16 * SUB;          // wave tile
  static constexpr int BT = 2 * WT;            // block tile
  static constexpr int BK = F::BK;
  static constexpr int P = F::P;
  static constexpr int VEC = 16 / ESZ;         // elements per 16-byte chunk
  static constexpr int ROW = BK + VEC;         // padded LDS row (elements)
  static constexpr int TILE = BT * ROW;        // elements of one piece tile
  static constexpr int CPR = BK / VEC;         // chunks per row
  static constexpr int CHUNKS = BT * CPR;
  static constexpr int LOADS = CHUNKS / 256;   // chunks per thread per piece tile
  static constexpr int STAGE = 2 * NPROD * P * TILE;  // X_I, Y_J (then Y_I, X_J)
  static constexpr int LDS_MAIN = 2 * STAGE * ESZ;
  static constexpr int HR = BT < 128 ? BT : 128;      // epilogue rows staged at once
  static constexpr int LDS_CTILE = HR * (BT + 1) * 4;
  static constexpr int LDS_BYTES = LDS_MAIN > LDS_CTILE ? LDS_MAIN : LDS_CTILE;
  static constexpr int SMEM_ELEMS = (LDS_BYTES + ESZ - 1) / ESZ;
  static_assert(CHUNKS % 256 == 0, "tile too small for 256 threads");
  static_assert(LDS_BYTES <= 163840, "LDS budget of one workgroup exceeded");
};

typedef float f32x16 __attribute__((ext_vector_type(16)));

// 32x32x16 MFMA: lane l holds row (column) l % 32 and k = 8 (l / 32) .. +7 of
// A (B); result element e of lane l is row 8 (e / 4) + 4 (l / 32) + e % 4,
// column l % 32.  One wave reaches the matrix-core peak with this shape,
// which it does not with 16x16x32.
__device__ __forceinline__ f32x16 mfma32(u32x4 a, u32x4 b, f32x16 c, int fmt) {
  if (fmt == FMT_BF16 || fmt == FMT_BF16X3 || fmt == FMT_BF16X9)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c, int fmt) {
  if (fmt == FMT_BF16 || fmt == FMT_BF16X3 || fmt == FMT_BF16X9)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// Epilogue for tile rows [r0, r0 + NR) staged in LDS (C holds local rows
// 0..NR-1, row stride BT + 1).  The primary block (i, j) is written along j,
// the mirror block (j, i) along i, so both are coalesced 16-byte stores.
// Tiles on the diagonal write only i <= j first and i < j in the mirror; a
// mirror store never lands on an element another phase reads (it is below
// the diagonal), so the phases of one tile do not interfere.
// cscale multiplies the staged values (the 2^-16 of FP16x3 products when the
// tile stored raw accumulators).
// Rows one epilogue_rows call covers at most (the LDS C tile holds 128 rows).
template <int BT>
constexpr int NR_MAX = BT < 128 ? BT : 128;

// A^3 quad of the polynomial epilogue at idx.  Products of binary16x3
// operands (A3PCS) read it from the pieces xh / xl of their X (the A^3 of the
// polynomial product), raw here and decoded by a3_value4 once loaded, so that
// the loads of several quads stay in flight together; the other formats read
// the float32 copy p3.  A compile-time choice: the other kernels' epilogues
// stay as they are.
template <bool A3PCS>
__device__ __forceinline__ f32x4 a3_load4(__attribute__((address_space(1))) const float* p3,
                                          __attribute__((address_space(1))) const uint16_t* xh,
                                          __attribute__((address_space(1))) const uint16_t* xl, int64_t idx) {
  if constexpr (A3PCS) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 h = *(const __attribute__((address_space(1))) u32x2*)(xh + idx);
    const u32x2 l = *(const __attribute__((address_space(1))) u32x2*)(xl + idx);
    // elements go through scalars: __builtin_bit_cast of a vector subscript
    // (v[i]) reads element 0 whatever i is
    const uint32_t h0 = h.x, h1 = h.y, l0 = l.x, l1 = l.y;
    return f32x4{__builtin_bit_cast(float, h0), __builtin_bit_cast(float, h1), __builtin_bit_cast(float, l0),
                 __builtin_bit_cast(float, l1)};
  } else {
    return *(const __attribute__((address_space(1))) f32x4*)(p3 + idx);
  }
}
template <bool A3PCS>
__device__ __forceinline__ f32x4 a3_value4(f32x4 raw) {
  if constexpr (A3PCS) {
    const float r0 = raw.x, r1 = raw.y, r2 = raw.z, r3 = raw.w;
    const uint32_t h0 = __builtin_bit_cast(uint32_t, r0), h1 = __builtin_bit_cast(uint32_t, r1);
    const uint32_t l0 = __builtin_bit_cast(uint32_t, r2), l1 = __builtin_bit_cast(uint32_t, r3);
    return f32x4{f16x3_value((uint16_t)h0, (uint16_t)l0), f16x3_value((uint16_t)(h0 >> 16), (uint16_t)(l0 >> 16)),
                 f16x3_value((uint16_t)h1, (uint16_t)l1), f16x3_value((uint16_t)(h1 >> 16), (uint16_t)(l1 >> 16))};
  } else {
    return raw;
  }
}

// Primary and mirror passes of a polynomial / store epilogue with the output
// operand format OF fixed at compile time, so that the per-element split is
// straight-line code and the loops stay compact (the instruction cache holds
// them; a runtime format switch unrolled per element did not fit).
template <int BT, int NT, int OF, int GMAX = 8, bool A3PCS = false>
__device__ __forceinline__ void epilogue_store_rows(float* C, int r0, int NR, float cscale, int row0, int col0,
                                                    bool diag, int64_t ld,
                                                    __attribute__((address_space(1))) float* c32,
                                                    __attribute__((address_space(1))) const float* p3,
                                                    __attribute__((address_space(1))) const uint16_t* xh,
                                                    __attribute__((address_space(1))) const uint16_t* xl,
                                                    uint16_t* const outp[3], bool poly, bool half_id, float ca,
                                                    float cb, float cc) {
  constexpr int CS = BT + 1;
  constexpr int Q4 = BT / 4;
  constexpr bool pieces = OF != FMT_F32;
  constexpr int NP = OF == FMT_F16X3 || OF == FMT_BF16X3 ? 2 : (OF == FMT_BF16X9 ? 3 : 1);
  const int tid = threadIdx.x;
  const int nq = NR / 4;
  __attribute__((address_space(1))) uint16_t* op[3] = {gbl(outp[0]), gbl(outp[1]), gbl(outp[2])};
  auto put4 = [&](int64_t idx, const float w[4]) {
    uint16_t pc[4][3];
#pragma unroll
    for (int e = 0; e < 4; ++e) split_value(OF, w[e], pc[e]);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
      const u32x2 pk = {(uint32_t)pc[0][q] | ((uint32_t)pc[1][q] << 16),
                        (uint32_t)pc[2][q] | ((uint32_t)pc[3][q] << 16)};
      *(__attribute__((address_space(1))) u32x2*)(op[q] + idx) = pk;
    }
  };
  auto put1 = [&](int64_t idx, float w) {
    uint16_t pc[3];
    split_value(OF, w, pc);
#pragma unroll
    for (int q = 0; q < NP; ++q) op[q][idx] = pc[q];
  };
  // pass 1: primary block, rows of the tile.  The polynomial's operands
  // (A and A^3 of this block) are loaded G quads ahead, so that a thread has
  // G loads in flight instead of waiting out each one before its stores.
  constexpr int G0 = (NR_MAX<BT> * Q4 / NT) < 8 ? (NR_MAX<BT> * Q4 / NT > 0 ? NR_MAX<BT> * Q4 / NT : 1) : 8;
  constexpr int G = G0 < GMAX ? G0 : GMAX;
#pragma unroll 1
  for (int q0 = tid; q0 < NR * Q4; q0 += NT * G) {
    f32x4 a4g[G], p4g[G];
    if (poly) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int q = q0 + g * NT;
        if (q < NR * Q4) {
          const int64_t idx = (int64_t)(row0 + q / Q4) * ld + col0 + (q % Q4) * 4;
          a4g[g] = *(const __attribute__((address_space(1))) f32x4*)(c32 + idx);
          p4g[g] = a3_load4<A3PCS>(p3, xh, xl, idx);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g * NT;
      if (q >= NR * Q4) break;
      const int r = q / Q4, c0 = (q % Q4) * 4;
      const int i = row0 + r, j0 = col0 + c0;
      const int64_t idx = (int64_t)i * ld + j0;
      float w[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = C[r * CS + c0 + e] * cscale;
      if (poly) {
        const f32x4 a4 = a4g[g];
        const f32x4 p4 = a3_value4<A3PCS>(p4g[g]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float x = ca * a4[e];
          x = fmaf(cb, p4[e], x);
          x = fmaf(cc, w[e], x);
          if (half_id) x = (x + (i == j0 + e ? 1.0f : 0.0f)) * 0.5f;
          w[e] = x;
        }
      }
      if (!diag) {
        if (c32) *(__attribute__((address_space(1))) f32x4*)(c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
        if constexpr (pieces) put4(idx, w);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (i <= j0 + e) {
            if (c32) c32[idx + e] = w[e];
            if constexpr (pieces) put1(idx + e, w[e]);
          }
        }
      }
      if (poly || cscale != 1.0f) {  // pass 2 reads final values
#pragma unroll
        for (int e = 0; e < 4; ++e) C[r * CS + c0 + e] = w[e];
      }
    }
  }
  __syncthreads();
  // pass 2: mirror block, columns of the tile
#pragma unroll 2
  for (int q = tid; q < BT * nq; q += NT) {
    const int c = q / nq, rr = (q % nq) * 4;
    const int j = col0 + c, i0 = row0 + rr;
    const int64_t idx = (int64_t)j * ld + i0;
    float w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = C[(rr + e) * CS + c];
    if (!diag) {
      if (c32) *(__attribute__((address_space(1))) f32x4*)(c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
      if constexpr (pieces) put4(idx, w);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (i0 + e < j) {
          if (c32) c32[idx + e] = w[e];
          if constexpr (pieces) put1(idx + e, w[e]);
        }
      }
    }
  }
}

// FMT_F16MX outputs: the same two passes, with each 32-value block of a
// stored row split together (mx_split4 over the eight lanes of its quads).
// Pass 1 first leaves the final values of the staged rows in C.  In a
// diagonal tile a block that contains the diagonal is stored partly by each
// pass (row i at j >= i from row i of C, at j < i from column i), so the
// staged rows' diagonal 32 x 32 blocks are made symmetric in C before any of
// them is stored: both passes then split the same 32 values and write the
// same scale.  Staged rows start at multiples of 32, so such a block is
// always staged whole.
__device__ __forceinline__ void mx_store1(uint16_t* const pieces[3], int64_t r, int c, int64_t ld, uint16_t hi,
                                          uint32_t h8, uint32_t l8, int e) {
  *gbl(pieces[0] + r * ld + c) = hi;
  char* b8 = (char*)pieces[1] + mx_byte(r, c, ld);
  *gbl(b8) = (char)(h8 >> (8 * e));
  *gbl(b8 + 32) = (char)(l8 >> (8 * e));
}

template <int BT, int NT, int GMAX = 8, bool A3PCS = false>
__device__ __forceinline__ void epilogue_store_rows_mx(float* C, int r0, int NR, float cscale, int row0, int col0,
                                                       bool diag, int64_t ld,
                                                       __attribute__((address_space(1))) float* c32,
                                                       __attribute__((address_space(1))) const float* p3,
                                                       __attribute__((address_space(1))) const uint16_t* xh,
                                                       __attribute__((address_space(1))) const uint16_t* xl,
                                                       uint16_t* const outp[3], bool poly, bool half_id, float ca,
                                                       float cb, float cc) {
  constexpr int CS = BT + 1;
  constexpr int Q4 = BT / 4;
  static_assert(Q4 % 8 == 0, "a 32-value block is 8 quads of one row");
  const int tid = threadIdx.x;
  const int nq = NR / 4;
  constexpr int G0 = (NR_MAX<BT> * Q4 / NT) < 8 ? (NR_MAX<BT> * Q4 / NT > 0 ? NR_MAX<BT> * Q4 / NT : 1) : 8;
  constexpr int G = G0 < GMAX ? G0 : GMAX;
  // pass 1: final values (and, off the diagonal, their stores)
#pragma unroll 1
  for (int q0 = tid; q0 < NR * Q4; q0 += NT * G) {
    f32x4 a4g[G], p4g[G];
    if (poly) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int q = q0 + g * NT;
        if (q < NR * Q4) {
          const int64_t idx = (int64_t)(row0 + q / Q4) * ld + col0 + (q % Q4) * 4;
          a4g[g] = *(const __attribute__((address_space(1))) f32x4*)(c32 + idx);
          p4g[g] = a3_load4<A3PCS>(p3, xh, xl, idx);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g * NT;
      if (q >= NR * Q4) break;  // uniform over each wave (NR * Q4 is a multiple of 64)
      const int r = q / Q4, c0 = (q % Q4) * 4;
      const int i = row0 + r, j0 = col0 + c0;
      float w[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = C[r * CS + c0 + e] * cscale;
      if (poly) {
        const f32x4 p4 = a3_value4<A3PCS>(p4g[g]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float x = ca * a4g[g][e];
          x = fmaf(cb, p4[e], x);
          x = fmaf(cc, w[e], x);
          if (half_id) x = (x + (i == j0 + e ? 1.0f : 0.0f)) * 0.5f;
          w[e] = x;
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) C[r * CS + c0 + e] = w[e];
      if (!diag) {
        const int64_t idx = (int64_t)i * ld + j0;
        if (c32) *(__attribute__((address_space(1))) f32x4*)(c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
        uint16_t hi[4];
        uint32_t h8, l8, sc;
        mx_split4(w, hi, h8, l8, sc);
        mx_store4(outp, i, j0, ld, hi, h8, l8, sc);
      }
    }
  }
  __syncthreads();
  if (diag) {
    // lower half of the staged rows' diagonal 32 x 32 blocks := upper half
    for (int e = tid; e < NR * 32; e += NT) {
      const int r = e >> 5, R = r0 + r, c = (R & ~31) + (e & 31);
      if (c < R) C[r * CS + c] = C[(c - r0) * CS + R];
    }
    __syncthreads();
#pragma unroll 2
    for (int q = tid; q < NR * Q4; q += NT) {
      const int r = q / Q4, c0 = (q % Q4) * 4;
      const int i = row0 + r, j0 = col0 + c0;
      float w[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = C[r * CS + c0 + e];
      uint16_t hi[4];
      uint32_t h8, l8, sc;
      mx_split4(w, hi, h8, l8, sc);
      const int64_t idx = (int64_t)i * ld + j0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (i <= j0 + e) {
          if (c32) c32[idx + e] = w[e];
          mx_store1(outp, i, j0 + e, ld, hi[e], h8, l8, e);
        }
      }
      if ((c0 & 31) == 0 && (j0 | 31) >= i) *gbl(outp[2] + mx_scale_index(i, j0, ld)) = (uint16_t)sc;
    }
  }
  // pass 2: mirror block, columns of the tile (stored row j, columns i0 ..)
#pragma unroll 2
  for (int q = tid; q < BT * nq; q += NT) {
    const int c = q / nq, rr = (q % nq) * 4;
    const int j = col0 + c, i0 = row0 + rr;
    float w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = C[(rr + e) * CS + c];
    uint16_t hi[4];
    uint32_t h8, l8, sc;
    mx_split4(w, hi, h8, l8, sc);
    const int64_t idx = (int64_t)j * ld + i0;
    if (!diag) {
      if (c32) *(__attribute__((address_space(1))) f32x4*)(c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
      mx_store4(outp, j, i0, ld, hi, h8, l8, sc);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (i0 + e < j) {
          if (c32) c32[idx + e] = w[e];
          mx_store1(outp, j, i0 + e, ld, hi[e], h8, l8, e);
        }
      }
      if ((rr & 31) == 0 && (i0 & ~31) < j) *gbl(outp[2] + mx_scale_index(j, i0, ld)) = (uint16_t)sc;
    }
  }
}

// MXOUT = false leaves out the FMT_F16MX store path (kernels that never write
// it: the float32 and averaged products, whose registers it would crowd).
// EPI_PARTIAL (split products): the staged rows (local rows 0..NR-1 of C,
// row stride BT + 1), unscaled, to rows r0 .. r0 + NR of the chunk's partial
// tile (row-major BT x BT floats).
template <int BT, int NT>
__device__ __forceinline__ void partial_rows(float* ptile, const float* C, int r0, int NR) {
  constexpr int CS = BT + 1;
  constexpr int Q4 = BT / 4;
  auto dst = gbl(ptile);
  for (int q = threadIdx.x; q < NR * Q4; q += NT) {
    const int r = q / Q4, c0 = (q % Q4) * 4;
    const float* src = C + r * CS + c0;
    *(__attribute__((address_space(1))) f32x4*)(dst + (int64_t)(r0 + r) * BT + c0) =
        f32x4{src[0], src[1], src[2], src[3]};
  }
}

template <int BT, int NT = 256, int GMAX = 8, bool MXOUT = true, bool A3PCS = false>
__device__ __forceinline__ void epilogue_rows(const ProdDesc& d, int bi, int bj, float* C, int r0, int NR,
                                              float cscale = 1.0f) {
  constexpr int CS = BT + 1;
  constexpr int Q4 = BT / 4;
  const int tid = threadIdx.x;
  const int row0 = bi * BT + r0, col0 = bj * BT;
  const bool diag = bi == bj;
  const int64_t ld = d.ld;
  const int nq = NR / 4;
  // Every descriptor field is copied to a local before the first store: the
  // descriptor may live in device memory (grouped launches), where the
  // compiler would otherwise reload it after each store it cannot prove
  // disjoint from it.
  auto c32 = gbl(d.c32);
  auto p3 = gbl(d.p3);  // A3PCS: A^3 from the pieces of X instead (binary16x3 products)
  auto xh = gbl((const uint16_t*)d.x[0]);
  auto xl = gbl((const uint16_t*)d.x[1]);
  const int mode = d.mode, out_fmt = d.out_fmt, n = d.n, out_f64 = d.out_f64;
  const float ca = d.a, cb = d.b, cc = d.c;
  uint16_t* const outp[3] = {d.outp[0], d.outp[1], d.outp[2]};
  if (mode == EPI_OUTPUT) {
    auto out64 = gbl((double*)d.out);
    auto out32 = gbl((float*)d.out);
    const double s = (d.scale ? *d.scale : 1.0) * (double)cscale;
    const int64_t old = d.out_ld;
    int* const flag = d.nonfinite;
    // (i, j) at i + j*ld: contiguous along i.  A diverged chain (spectrum
    // outside [-1, 1] without the bound) shows up here as Inf/NaN: the
    // first pass covers every written value once (the mirror pass repeats
    // them), so it alone checks them and raises the caller's flag.
    bool bad = false;
#pragma unroll 2
    for (int q = tid; q < BT * nq; q += NT) {
      const int c = q / nq, rr = (q % nq) * 4;
      const int j = col0 + c;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = row0 + rr + e;
        if (i < n && j < n && (!diag || i <= j)) {
          const double v = (double)C[(rr + e) * CS + c] * s;
          bad |= !isfinite(v);
          if (out_f64) out64[(int64_t)i + (int64_t)j * old] = v;
          else out32[(int64_t)i + (int64_t)j * old] = (float)v;
        }
      }
    }
    if (bad && flag) *gbl(flag) = 1;
    // (j, i) at j + i*ld: contiguous along j
#pragma unroll 2
    for (int q = tid; q < NR * Q4; q += NT) {
      const int r = q / Q4, c0 = (q % Q4) * 4;
      const int i = row0 + r;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = col0 + c0 + e;
        if (i < n && j < n && (!diag || i < j)) {
          const double v = (double)C[r * CS + c0 + e] * s;
          if (out_f64) out64[(int64_t)j + (int64_t)i * old] = v;
          else out32[(int64_t)j + (int64_t)i * old] = (float)v;
        }
      }
    }
    return;
  }
  const bool poly = mode == EPI_POLY || mode == EPI_POLY_HALF_ID;
  const bool half_id = mode == EPI_POLY_HALF_ID;
#define PSDP_EPI(OF)                                                                                      \
  epilogue_store_rows<BT, NT, OF, GMAX, A3PCS>(C, r0, NR, cscale, row0, col0, diag, ld, c32, p3, xh, xl, outp,  \
                                              poly, half_id, ca, cb, cc)
  if constexpr (MXOUT) {
    if (out_fmt == FMT_F16MX) {
      epilogue_store_rows_mx<BT, NT, GMAX, A3PCS>(C, r0, NR, cscale, row0, col0, diag, ld, c32, p3, xh, xl, outp,
                                                  poly, half_id, ca, cb, cc);
      return;
    }
  }
  switch (out_fmt) {
    case FMT_F16: PSDP_EPI(FMT_F16); break;
    case FMT_F16X3: PSDP_EPI(FMT_F16X3); break;
    case FMT_BF16: PSDP_EPI(FMT_BF16); break;
    case FMT_BF16X3: PSDP_EPI(FMT_BF16X3); break;
    case FMT_BF16X9: PSDP_EPI(FMT_BF16X9); break;
    case FMT_E4M3: PSDP_EPI(FMT_E4M3); break;
    default: PSDP_EPI(FMT_F32); break;
  }
#undef PSDP_EPI
}

// MFMAs over one BK-wide LDS slab: acc += X-rows x Y-rows^T.
template <int FMT, int SUB>
__device__ __forceinline__ void mma_slab(const typename Tiling<FMT, SUB>::E* X,
                                         const typename Tiling<FMT, SUB>::E* Y, int arow, int brow, int lane,
                                         f32x4 (&acc)[SUB][SUB]) {
  using TL = Tiling<FMT, SUB>;
  constexpr int P = TL::P, BK = TL::BK, ROW = TL::ROW, TILE = TL::TILE;
  if constexpr (FMT == FMT_F32) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      f32x4 af[SUB], bf[SUB];
      const int ko = kk + (lane >> 4) * 4;
#pragma unroll
      for (int s = 0; s < SUB; ++s) {
        af[s] = *(const f32x4*)(X + arow + s * 16 * ROW + ko);
        bf[s] = *(const f32x4*)(Y + brow + s * 16 * ROW + ko);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int si = 0; si < SUB; ++si)
#pragma unroll
          for (int sj = 0; sj < SUB; ++sj)
            acc[si][sj] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[si][q], bf[sj][q], acc[si][sj], 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      // A fragments of all SUB row blocks stay live; B fragments are read one
      // column block at a time (same LDS reads, fewer live registers).
      u32x4 af[P][SUB];
      const int ko = kk + (lane >> 4) * 8;
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int s = 0; s < SUB; ++s) af[p][s] = *(const u32x4*)(X + p * TILE + arow + s * 16 * ROW + ko);
#pragma unroll
      for (int sj = 0; sj < SUB; ++sj) {
        u32x4 bf[P];
#pragma unroll
        for (int p = 0; p < P; ++p) bf[p] = *(const u32x4*)(Y + p * TILE + brow + sj * 16 * ROW + ko);
#pragma unroll
        for (int si = 0; si < SUB; ++si) {
          f32x4 c = acc[si][sj];
          if constexpr (P == 2) {  // x3: lo*hi, hi*lo, then hi*hi
            c = mfma16(af[1][si], bf[0], c, FMT);
            c = mfma16(af[0][si], bf[1], c, FMT);
          } else if constexpr (P == 3) {  // x9: smallest terms first
            c = mfma16(af[2][si], bf[2], c, FMT);
            c = mfma16(af[1][si], bf[2], c, FMT);
            c = mfma16(af[2][si], bf[1], c, FMT);
            c = mfma16(af[1][si], bf[1], c, FMT);
            c = mfma16(af[0][si], bf[2], c, FMT);
            c = mfma16(af[2][si], bf[0], c, FMT);
            c = mfma16(af[0][si], bf[1], c, FMT);
            c = mfma16(af[1][si], bf[0], c, FMT);
          }
          acc[si][sj] = mfma16(af[0][si], bf[0], c, FMT);
        }
      }
    }
  }
}

// One upper-triangle tile (bi <= bj) of C = X Y.  AVG: also accumulate
// C^T on the same tile (rows of Y for bi, rows of X for bj) and store the
// average (C + C^T)/2, i.e. the symmetric (Jordan) product.
template <int FMT, int SUB, bool AVG>
__device__ __forceinline__ void sym_tile(const ProdDesc& d, int bi, int bj,
                                         typename Tiling<FMT, SUB, AVG>::E* smem) {
  using TL = Tiling<FMT, SUB, AVG>;
  using E = typename TL::E;
  constexpr int P = TL::P, LOADS = TL::LOADS, BK = TL::BK, ROW = TL::ROW, TILE = TL::TILE;
  constexpr int NP = TL::NPROD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t ld = d.ld;
  const int row0 = bi * TL::BT, col0 = bj * TL::BT;

  // source panels: [0] X rows of bi, [1] Y rows of bj, [2] Y rows of bi, [3] X rows of bj
  const E* src[2 * NP][P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    src[0][p] = (const E*)d.x[p] + (int64_t)row0 * ld;
    src[1][p] = (const E*)d.y[p] + (int64_t)col0 * ld;
    if constexpr (AVG) {
      src[2][p] = (const E*)d.y[p] + (int64_t)row0 * ld;
      src[3][p] = (const E*)d.x[p] + (int64_t)col0 * ld;
    }
  }
  int goff[LOADS], soff[LOADS];
#pragma unroll
  for (int i = 0; i < LOADS; ++i) {
    const int c = tid + i * 256;
    const int r = c / TL::CPR, cc = (c % TL::CPR) * TL::VEC;
    goff[i] = r * (int)ld + cc;  // r < BT, so r*ld stays far below 2^31
    soff[i] = r * ROW + cc;
  }

  f32x4 acc[NP][SUB][SUB];
#pragma unroll
  for (int u = 0; u < NP; ++u)
#pragma unroll
    for (int a = 0; a < SUB; ++a)
#pragma unroll
      for (int b = 0; b < SUB; ++b) acc[u][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int arow = (wr * TL::WT + (lane & 15)) * ROW;
  const int brow = (wc * TL::WT + (lane & 15)) * ROW;
  const int KT = d.npad / BK;

  // Global -> registers -> LDS staging, double buffered: the next slab is
  // fetched into registers while the MFMAs consume the current LDS buffer.
  // The last slab is peeled so that loads and stores are unconditional.
  u32x4 rg[2 * NP][P][LOADS];
#define PSDP_GLOAD(k0)                                                          \
  _Pragma("unroll") for (int g = 0; g < 2 * NP; ++g)                            \
  _Pragma("unroll") for (int p = 0; p < P; ++p)                                 \
  _Pragma("unroll") for (int i = 0; i < LOADS; ++i)                             \
    rg[g][p][i] = *(const u32x4*)(src[g][p] + goff[i] + (k0));
#define PSDP_LSTORE(buf)                                                        \
  _Pragma("unroll") for (int g = 0; g < 2 * NP; ++g)                            \
  _Pragma("unroll") for (int p = 0; p < P; ++p)                                 \
  _Pragma("unroll") for (int i = 0; i < LOADS; ++i)                             \
    *(u32x4*)(smem + (buf) * TL::STAGE + (g * P + p) * TILE + soff[i]) = rg[g][p][i];
  // FP32 operands: every 16x16x4 MFMA rounds the running sum once, i.e. n/4
  // roundings at the sum's full magnitude per entry (the 16-bit MFMAs add 16
  // products per rounding).  Two-level accumulation: FLUSH slabs (128 k) go
  // into a zeroed partial sum that is then added to the total, so an entry
  // sees 32 roundings at the partial's magnitude plus n/128 at its own
  // (measured rms error of the float32 product at n = 4096 was 2.8x that of
  // the 16-bit formats on exactly representable data before this).
  constexpr bool TWO = FMT == FMT_F32 && !AVG;
  constexpr int FLUSH = 4;
  f32x4 part[SUB][SUB];
  auto clear_part = [&]() {
#pragma unroll
    for (int a = 0; a < SUB; ++a)
#pragma unroll
      for (int b = 0; b < SUB; ++b) part[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  };
  auto flush_part = [&]() {
#pragma unroll
    for (int a = 0; a < SUB; ++a)
#pragma unroll
      for (int b = 0; b < SUB; ++b) {
        acc[0][a][b] += part[a][b];
        part[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
  };
  if constexpr (TWO) clear_part();
#define PSDP_MMA(buf)                                                                          \
  {                                                                                            \
    const E* base_ = smem + (buf) * TL::STAGE;                                                 \
    if constexpr (TWO) {                                                                       \
      mma_slab<FMT, SUB>(base_, base_ + P * TILE, arow, brow, lane, part);                     \
    } else {                                                                                   \
      mma_slab<FMT, SUB>(base_, base_ + P * TILE, arow, brow, lane, acc[0]);                   \
    }                                                                                          \
    if constexpr (AVG)                                                                         \
      mma_slab<FMT, SUB>(base_ + 2 * P * TILE, base_ + 3 * P * TILE, arow, brow, lane, acc[NP - 1]); \
  }
  PSDP_GLOAD(0)
  PSDP_LSTORE(0)
  __syncthreads();
  for (int kt = 0; kt < KT - 1; ++kt) {
    PSDP_GLOAD((kt + 1) * BK)
    PSDP_MMA(kt & 1)
    if constexpr (TWO) {
      if ((kt + 1) % FLUSH == 0) flush_part();
    }
    PSDP_LSTORE((kt + 1) & 1)
    __syncthreads();
  }
  PSDP_MMA((KT - 1) & 1)
  if constexpr (TWO) flush_part();
  __syncthreads();
#undef PSDP_GLOAD
#undef PSDP_LSTORE
#undef PSDP_MMA

  // ---------------------------------------------------------------- epilogue
  // Accumulators go to LDS with compile-time indices only (they stay in
  // registers), HR rows at a time, then epilogue_rows writes them out.
  constexpr float PSCALE = (FMT == FMT_F16X3 || FMT == FMT_E4M3) ? F16X3_PROD_SCALE : 1.0f;
  float* Ct = (float*)smem;
  constexpr int CS = TL::BT + 1;
  constexpr int PHASES = TL::BT / TL::HR;
#pragma unroll
  for (int ph = 0; ph < PHASES; ++ph) {
    // with two phases, wave row wr owns tile rows [wr*WT, wr*WT + WT) = phase wr
    if (PHASES == 1 || wr == ph) {
#pragma unroll
      for (int si = 0; si < SUB; ++si)
#pragma unroll
        for (int sj = 0; sj < SUB; ++sj)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float v = acc[0][si][sj][e];
            if constexpr (AVG) v = 0.5f * (v + acc[NP - 1][si][sj][e]);
            const int r = wr * TL::WT + si * 16 + (lane >> 4) * 4 + e - ph * TL::HR;
            const int c = wc * TL::WT + sj * 16 + (lane & 15);
            Ct[r * CS + c] = v * PSCALE;
          }
    }
    __syncthreads();
    epilogue_rows<TL::BT, 256, 8, false, FMT == FMT_F16X3>(d, bi, bj, Ct, ph * TL::HR, TL::HR);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS pipeline (16-bit formats).  Operand slabs go from global
// memory straight into LDS (global_load_lds_dwordx4: lane L of a wave writes
// bytes [16 L, 16 L + 16) of a 1 KB block whose base is wave-uniform), S
// stages deep, so that S - 1 slabs are in flight while one is consumed; no
// staging registers.  Slab rows are unpadded (RB = 64 or 128 bytes) and
// their 16-byte chunks are XOR swizzled: logical chunk c of row r sits at
// physical chunk c ^ swz(r), swz(r) = (r >> 2) & 3 for 64-byte rows and
// (r >> 1) & 7 for 128-byte rows.  A 32x32x16 fragment read (lane l: row
// l % 32, chunk 2 ks + l / 32) is then free of LDS bank conflicts; both swizzles were
// chosen by measurement (SQ_LDS_BANK_CONFLICT over candidate patterns),
// several XOR patterns that look conflict-free under a plain 64-bank model
// cost two cycles per read here.  Each global->LDS load reads whole 64- or
// 128-byte row segments with consecutive lanes, as coalescing wants.
// The loop waits for its own loads with an explicit vmcnt and crosses a bare
// s_barrier (a __syncthreads would also wait for the slabs still in flight).
template <int FMT, int SUB, bool WS = false>
struct DmaTiling {
  static constexpr int P = FmtTraits<FMT>::P;
  static constexpr int WT = 16 * SUB;
  static constexpr int BT = 2 * WT;
  // Small tiles keep LDS under 80 KB so that two workgroups share a CU.
  static constexpr int LDS_BUDGET = SUB <= 2 ? 81920 : 163840;
  // 64-element slabs make every row segment a whole 128-byte line (half as
  // many L2 requests as 32-element slabs) where two such stages fit in LDS;
  // with producer waves two-piece operands also take 32-element slabs (the
  // fragments of 64 would not fit 256 registers per wave).
  // FMT_F16MX stages the block scales of both panels after each stage's
  // operand bytes: 2 bytes per row and 32-k block.
  static constexpr bool MX = FMT == FMT_F16MX;
  static constexpr bool FIT64 = 2 * (2 * P * BT * 128 + (MX ? 8 * BT : 0)) <= LDS_BUDGET;
  static constexpr int BK = (FIT64 && !(WS && P == 2)) ? 64 : 32;  // slab width (elements)
  static constexpr int RB = BK * 2;                        // row bytes
  static constexpr int TILE_BYTES = BT * RB;               // one (panel, piece)
  static constexpr int STAGE_BYTES = 2 * P * TILE_BYTES;   // X rows of bi, Y rows of bj
  static constexpr int SCALE_BYTES = MX ? 2 * 2 * BT * (BK / 32) : 0;
  static constexpr int STAGE_STRIDE = STAGE_BYTES + SCALE_BYTES;
  static constexpr int SFIT = LDS_BUDGET / STAGE_STRIDE;
  static constexpr int S = SFIT >= 4 ? 4 : SFIT;           // pipeline depth (>= 2)
  static constexpr int L = STAGE_BYTES / 4096;             // loads per thread per stage
  static constexpr int BLOCKS = TILE_BYTES / 1024;         // 1 KB blocks per tile
  static constexpr int HR = BT < 128 ? BT : 128;
  static constexpr int LDS_MAIN = S * STAGE_STRIDE;
  static constexpr int LDS_CTILE = HR * (BT + 1) * 4;
  static constexpr int LDS_BYTES = LDS_MAIN > LDS_CTILE ? LDS_MAIN : LDS_CTILE;
  static_assert(STAGE_BYTES % 4096 == 0, "stage must be a whole number of 4 KB loads");
  static_assert(RB == 64 || RB == 128, "swizzle defined for 64- and 128-byte rows");
  static_assert(S >= 2 && S <= SFIT && LDS_BYTES <= LDS_BUDGET, "pipeline does not fit in LDS");
  static_assert((S - 1) * L <= 63, "vmcnt range");
};

// Chunk swizzle of slab row r (see the layout note above); depends on r mod 32
// only, so the 32-row blocks of one wave share it.
template <int RB>
__device__ __forceinline__ int dma_swz(int r) {
  if constexpr (RB == 64) return (r >> 2) & 3;
  else return (r >> 1) & 7;
}

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N <= 63, "vmcnt range");
  // s_waitcnt: vmcnt[3:0] | expcnt[6:4] = 7 | lgkmcnt[11:8] = 15 | vmcnt[5:4] at [15:14]
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Slabs [k0, k0 + KT) of a tile: the whole K, or for a split product
// (EPI_PARTIAL) chunk s of d.ksplit, each d.kchunk64 * 64 elements long (the
// host sizes the chunks so that none is empty).
template <int BK>
__device__ __forceinline__ int tile_k_range(const ProdDesc& d, int s, int& k0) {
  const int KT = d.npad / BK;
  if (d.mode != EPI_PARTIAL) {
    k0 = 0;
    return KT;
  }
  constexpr int PER = 64 / BK;  // slabs per 64 k
  k0 = s * d.kchunk64 * PER;
  return min(KT - k0, d.kchunk64 * PER);
}

// WS (producer waves): the workgroup has 8 waves; waves 0-3 compute (one per
// SIMD) and never issue a global->LDS load, waves 4-7 only issue them and
// wait for them.  A wave that issues a 1 KB global_load_lds stalls its own
// instruction stream for many cycles, and with one compute wave per SIMD
// those cycles are lost to the matrix cores; with a producer wave beside it
// the compute wave keeps issuing MFMAs.  Requires <= 256 registers per wave.
// PART: the chunk of a split product (EPI_PARTIAL) whose epilogue only copies
// the raw tile out (a separate instantiation, so that the register budget of
// the full epilogue is untouched).
template <int FMT, int SUB, bool WS, bool PART = false>
__device__ __forceinline__ void sym_tile_dma(const ProdDesc& d, int bi, int bj, char* smem, int split_slot = 0,
                                             int split_chunk = 0) {
  using DT = DmaTiling<FMT, SUB, WS>;
  constexpr int P = DT::P, BK = DT::BK, RB = DT::RB, S = DT::S, L = DT::L;
  constexpr int TB = DT::TILE_BYTES, SB = DT::STAGE_BYTES;
  constexpr int NT = WS ? 512 : 256;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform (SGPR)
  const bool compute = !WS || wave < 4;                         // issues MFMAs
  const bool loader = !WS || wave >= 4;                         // issues global->LDS loads
  const int dwave = WS ? wave - 4 : wave;                       // index among the loaders
  const int wr = (wave & 3) >> 1, wc = wave & 1;
  const int64_t ld = d.ld;
  const int row0 = bi * DT::BT, col0 = bj * DT::BT;

  // This thread's L loads per stage.  Load i of a wave fills one 1 KB block
  // (RPB whole rows) of one (panel, piece); its base pointer is wave-uniform
  // (scalar registers).  Lane l takes row l / CPR of the block and physical
  // chunk l % CPR, i.e. logical chunk (l % CPR) ^ swz(row).
  constexpr int CPR = RB / 16, RPB = 1024 / RB;
  const uint16_t* gsrc[L];
  int lofs[L], loff[L];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int q = dwave * L + i;
    const int gp = q / DT::BLOCKS, blk = q % DT::BLOCKS;
    const int g = gp / P, p = gp % P;
    // select chain rather than a dynamic index: keeps a descriptor held in
    // registers out of private memory
    const void* bx = p == 0 ? d.x[0] : (p == 1 ? d.x[1] : d.x[2]);
    const void* by = p == 0 ? d.y[0] : (p == 1 ? d.y[1] : d.y[2]);
    const void* base = g == 0 ? bx : by;
    gsrc[i] = (const uint16_t*)base + (int64_t)((g == 0 ? row0 : col0) + blk * RPB) * ld;
    lofs[i] = gp * TB + blk * 1024;
    const int r = blk * RPB + lane / CPR;
    loff[i] = (lane / CPR) * (int)ld + ((lane % CPR) ^ dma_swz<RB>(r)) * 8;
  }
#define PSDP_DMA_ISSUE(K0, BUF)                                                                      \
  _Pragma("unroll") for (int i = 0; i < L; ++i)                                                     \
      __builtin_amdgcn_global_load_lds((const void*)(gsrc[i] + (K0) + loff[i]),                     \
                                       (lds_ptr_t)(smem + (BUF) * SB + lofs[i]), 16, 0, 0);

  // Each wave: M x M blocks of 32 x 32 (M = SUB / 2), KS k-steps of 16 per slab.
  constexpr int M = SUB / 2, KS = BK / 16;
  static_assert(SUB % 2 == 0, "32-row blocks");
  f32x16 acc[M][M];
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.f;

  // the lane's row within this wave's A and B row blocks
  const int arow = (wr * DT::WT + (lane & 31)) * RB;
  const int brow = (wc * DT::WT + (lane & 31)) * RB;
  const int swzl = dma_swz<RB>(lane & 31);
  int KT = d.npad / BK, k0 = 0;  // slabs k0 .. k0 + KT (a chunk of K for a split product)
  if constexpr (PART) KT = tile_k_range<BK>(d, split_chunk, k0);

  // fa/fb hold the slab being multiplied; na receives the A fragments of the
  // next slab, and the B fragments of the next slab overwrite fb column block
  // by column block as soon as that block's MFMAs have been issued.
  u32x4 fa[KS][P][M], fb[KS][P][M], na[KS][P][M];
  auto frag_a = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + p * TB + arow + m * 32 * RB + 16 * ((ks * 2 + (lane >> 5)) ^ swzl));
  };
  auto frag_b = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + (P + p) * TB + brow + m * 32 * RB + 16 * ((ks * 2 + (lane >> 5)) ^ swzl));
  };

  // Pipeline: S slab buffers.  Iteration kt computes slab kt from registers
  // while slab kt + 1 is read from LDS and slabs kt + 2 .. kt + S are in
  // flight from global memory; the DMA for slab kt + S reuses the buffer of
  // slab kt, whose fragments every wave holds in registers by then.
  // Every iteration issues one stage (past the end: a copy of the last slab
  // into a buffer that is not read again), so that the wait counts are
  // uniform and the loop body is one block that the scheduler can interleave.
  if (loader) {
#pragma unroll
    for (int st = 0; st < S; ++st) { PSDP_DMA_ISSUE((k0 + (st < KT ? st : KT - 1)) * BK, st) }
    wait_vmcnt<(S - 1) * L>();
  }
  lds_barrier();
  if (compute) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int m = 0; m < M; ++m) {
          fa[ks][p][m] = frag_a(smem, ks, p, m);
          fb[ks][p][m] = frag_b(smem, ks, p, m);
        }
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  }

  // Three-piece (x9) products add 9 MFMAs per k-step into one sum, each
  // rounding it at full magnitude; they accumulate every 4 slabs (128 k) into
  // a zeroed partial sum that is then added to the total (two-level
  // accumulation: n/128 roundings at the total's magnitude instead of 9 n/16).
  constexpr bool TWO = P == 3;
  constexpr int FLUSH = 4;  // slabs per partial sum
  static_assert(!TWO || BK == 32, "x9 partial sums cover 128 k (same order for every tile size)");
  f32x16 part[M][M];
  if constexpr (TWO) {
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
      for (int b = 0; b < M; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) part[a][b][e] = 0.f;
  }
  auto flush_part = [&]() {
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
      for (int b = 0; b < M; ++b) {
        acc[a][b] += part[a][b];
#pragma unroll
        for (int e = 0; e < 16; ++e) part[a][b][e] = 0.f;
      }
  };

  // MFMAs of column block mj of k-step ks
#define PSDP_COLUMN_MMA(ks, mj)                                                                   \
  _Pragma("unroll") for (int mi = 0; mi < M; ++mi) {                                              \
    if constexpr (P == 3) { /* x9: smallest terms first, into the slab's partial sum */           \
      f32x16 c = part[mi][mj];                                                                    \
      c = mfma32(fa[ks][2][mi], fb[ks][2][mj], c, FMT);                                           \
      c = mfma32(fa[ks][1][mi], fb[ks][2][mj], c, FMT);                                           \
      c = mfma32(fa[ks][2][mi], fb[ks][1][mj], c, FMT);                                           \
      c = mfma32(fa[ks][1][mi], fb[ks][1][mj], c, FMT);                                           \
      c = mfma32(fa[ks][0][mi], fb[ks][2][mj], c, FMT);                                           \
      c = mfma32(fa[ks][2][mi], fb[ks][0][mj], c, FMT);                                           \
      c = mfma32(fa[ks][0][mi], fb[ks][1][mj], c, FMT);                                           \
      c = mfma32(fa[ks][1][mi], fb[ks][0][mj], c, FMT);                                           \
      part[mi][mj] = mfma32(fa[ks][0][mi], fb[ks][0][mj], c, FMT);                                \
    } else {                                                                                      \
      f32x16 c = acc[mi][mj];                                                                     \
      if constexpr (P == 2) { /* x3: lo*hi, hi*lo, then hi*hi */                                  \
        c = mfma32(fa[ks][1][mi], fb[ks][0][mj], c, FMT);                                         \
        c = mfma32(fa[ks][0][mi], fb[ks][1][mj], c, FMT);                                         \
      }                                                                                           \
      acc[mi][mj] = mfma32(fa[ks][0][mi], fb[ks][0][mj], c, FMT);                                 \
    }                                                                                             \
  }

  constexpr int NREAD = 2 * KS * P * M;  // LDS reads per slab
  constexpr int NMFMA = KS * M * M * (P == 1 ? 1 : (P == 2 ? 3 : 9));
  constexpr int MR = NMFMA / NREAD > 0 ? NMFMA / NREAD : 1;
  for (int kt = 0; kt + 1 < KT; ++kt) {
    // own reads of buffer kt % S complete, own DMA of slab kt + 1 landed
    if (loader) wait_vmcnt<(S - 2) * L>();
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    lds_barrier();
    if (loader) { PSDP_DMA_ISSUE((k0 + (kt + S < KT ? kt + S : KT - 1)) * BK, kt % S) }
    if (!compute) continue;
    const char* nxt = smem + ((kt + 1) % S) * SB;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int m = 0; m < M; ++m) na[ks][p][m] = frag_a(nxt, ks, p, m);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int mj = 0; mj < M; ++mj) {
        PSDP_COLUMN_MMA(ks, mj)
#pragma unroll
        for (int p = 0; p < P; ++p) fb[ks][p][mj] = frag_b(nxt, ks, p, mj);
      }
    // Interleave global->LDS loads and LDS reads with the MFMAs so that the
    // memory pipes work while the matrix cores do (issuing all reads first
    // stalls the wave on the LDS command queue with the matrix cores idle).
#pragma unroll
    for (int g = 0; g < NREAD; ++g) {
      if (!WS && g < L) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      __builtin_amdgcn_sched_group_barrier(0x008, MR, 0);  // MFMA
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (TWO) {
      if ((kt + 1) % FLUSH == 0) flush_part();
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int m = 0; m < M; ++m) fa[ks][p][m] = na[ks][p][m];
  }
  if (compute) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int mj = 0; mj < M; ++mj) { PSDP_COLUMN_MMA(ks, mj) }
    if constexpr (TWO) flush_part();
  }
#undef PSDP_COLUMN_MMA
  __builtin_amdgcn_sched_barrier(0);  // keep the epilogue's accumulator reads after the MFMAs
  if (loader) wait_vmcnt<0>();  // trailing stage copies land before the epilogue reuses LDS
#undef PSDP_DMA_ISSUE
  lds_barrier();  // every wave is done reading the stages before they hold C
  float* ptile = nullptr;  // split products: this chunk's partial tile
  if constexpr (PART) ptile = d.part + ((int64_t)split_slot * d.ksplit + split_chunk) * 65536;

  // ---------------------------------------------------------------- epilogue
  constexpr float PSCALE = (FMT == FMT_F16X3 || FMT == FMT_E4M3) ? F16X3_PROD_SCALE : 1.0f;
  float* Ct = (float*)smem;
  constexpr int CS = DT::BT + 1;
  constexpr int PHASES = DT::BT / DT::HR;
#pragma unroll
  for (int ph = 0; ph < PHASES; ++ph) {
    if (compute && (PHASES == 1 || wr == ph)) {
#pragma unroll
      for (int mi = 0; mi < M; ++mi)
#pragma unroll
        for (int mj = 0; mj < M; ++mj)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int r = wr * DT::WT + mi * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3) - ph * DT::HR;
            const int c = wc * DT::WT + mj * 32 + (lane & 31);
            Ct[r * CS + c] = acc[mi][mj][e];
          }
    }
    __syncthreads();
    if constexpr (PART) partial_rows<DT::BT, NT>(ptile, Ct, ph * DT::HR, DT::HR);
    else epilogue_rows<DT::BT, NT, 8, true, FMT == FMT_F16X3>(d, bi, bj, Ct, ph * DT::HR, DT::HR, PSCALE);
    __syncthreads();
  }
}

template <int FMT, int SUB, bool WS, bool PART = false>
__global__ __launch_bounds__(WS ? 512 : 256) void sym_prod_dma_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT, SUB, WS>::LDS_BYTES];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile_dma<FMT, SUB, WS, PART>(d, t.bi, t.bj, smem, t.mat, t.pad);
}

// 256 x 256 tile with 8 waves (two per SIMD): wave w owns rows wr*128 .. +128
// (wr = w >> 2) and columns wc*64 .. +64 (wc = w & 3), 4 x 2 blocks of 32 x 32
// in 128 accumulator registers, so two waves share a SIMD and one wave's load
// issue and barrier waits overlap the other's MFMAs.  Same slabs, stages,
// swizzles and per-entry k order as sym_tile_dma<FMT, 8> (bitwise the same
// results); every wave issues its share of the global->LDS loads.  The
// fragments of one slab (all k-steps) live in registers; those of the next
// slab replace them k-step by k-step once that k-step's MFMAs are issued.
// The epilogue preloads at most 2 quads per thread (256 registers per wave).
template <int FMT>
__device__ __forceinline__ void sym_tile_dma8(const ProdDesc& d, const TileRef& t, char* smem) {
  const int bi = t.bi, bj = t.bj;
  using DT = DmaTiling<FMT, 8, false>;
  constexpr int P = DT::P, BK = DT::BK, RB = DT::RB, S = DT::S;
  constexpr int TB = DT::TILE_BYTES, SB = DT::STAGE_BYTES;
  constexpr int NT = 512;
  constexpr int L = SB / (NT * 16);  // 1 KB loads per thread per stage
  static_assert(SB % (NT * 16) == 0 && L >= 1, "stage must split over 8 waves");
  static_assert((S - 1) * L <= 63, "vmcnt range");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 2, wc = wave & 3;
  const int64_t ld = d.ld;
  const int row0 = bi * 256, col0 = bj * 256;

  constexpr int CPR = RB / 16, RPB = 1024 / RB;
  const uint16_t* gsrc[L];
  int lofs[L], loff[L];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int q = wave * L + i;
    const int gp = q / DT::BLOCKS, blk = q % DT::BLOCKS;
    const int g = gp / P, p = gp % P;
    const void* bx = p == 0 ? d.x[0] : (p == 1 ? d.x[1] : d.x[2]);
    const void* by = p == 0 ? d.y[0] : (p == 1 ? d.y[1] : d.y[2]);
    const void* base = g == 0 ? bx : by;
    gsrc[i] = (const uint16_t*)base + (int64_t)((g == 0 ? row0 : col0) + blk * RPB) * ld;
    lofs[i] = gp * TB + blk * 1024;
    const int r = blk * RPB + lane / CPR;
    loff[i] = (lane / CPR) * (int)ld + ((lane % CPR) ^ dma_swz<RB>(r)) * 8;
  }
#define PSDP_DMA8_ISSUE(K0, BUF)                                                                     \
  _Pragma("unroll") for (int i = 0; i < L; ++i)                                                     \
      __builtin_amdgcn_global_load_lds((const void*)(gsrc[i] + (K0) + loff[i]),                     \
                                       (lds_ptr_t)(smem + (BUF) * SB + lofs[i]), 16, 0, 0);

  constexpr int MA = 4, MB = 2, KS = BK / 16;
  f32x16 acc[MA][MB];
#pragma unroll
  for (int a = 0; a < MA; ++a)
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.f;

  const int arow = (wr * 128 + (lane & 31)) * RB;
  const int brow = (wc * 64 + (lane & 31)) * RB;
  const int swzl = dma_swz<RB>(lane & 31);
  int k0;
  const int KT = tile_k_range<BK>(d, t.pad, k0);
#pragma unroll
  for (int i = 0; i < L; ++i) gsrc[i] += (int64_t)k0 * BK;

  u32x4 fa[KS][P][MA], fb[KS][P][MB];
  auto frag_a = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + p * TB + arow + m * 32 * RB + 16 * ((ks * 2 + (lane >> 5)) ^ swzl));
  };
  auto frag_b = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + (P + p) * TB + brow + m * 32 * RB + 16 * ((ks * 2 + (lane >> 5)) ^ swzl));
  };

#pragma unroll
  for (int st = 0; st < S; ++st) { PSDP_DMA8_ISSUE((st < KT ? st : KT - 1) * BK, st) }
  wait_vmcnt<(S - 1) * L>();
  lds_barrier();
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int m = 0; m < MA; ++m) fa[ks][p][m] = frag_a(smem, ks, p, m);
#pragma unroll
      for (int m = 0; m < MB; ++m) fb[ks][p][m] = frag_b(smem, ks, p, m);
    }
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)

#define PSDP_KSTEP_MMA(ks)                                                                        \
  _Pragma("unroll") for (int mj = 0; mj < MB; ++mj)                                               \
  _Pragma("unroll") for (int mi = 0; mi < MA; ++mi) {                                             \
    f32x16 c = acc[mi][mj];                                                                       \
    if constexpr (P == 2) { /* x3: lo*hi, hi*lo, then hi*hi */                                    \
      c = mfma32(fa[ks][1][mi], fb[ks][0][mj], c, FMT);                                           \
      c = mfma32(fa[ks][0][mi], fb[ks][1][mj], c, FMT);                                           \
    }                                                                                             \
    acc[mi][mj] = mfma32(fa[ks][0][mi], fb[ks][0][mj], c, FMT);                                   \
  }

  constexpr int NREAD = KS * P * (MA + MB);
  constexpr int NMFMA = KS * MA * MB * (P == 1 ? 1 : 3);
  constexpr int MR = NMFMA / NREAD > 0 ? NMFMA / NREAD : 1;
  for (int kt = 0; kt + 1 < KT; ++kt) {
    wait_vmcnt<(S - 2) * L>();          // own DMA of slab kt + 1 landed
    __builtin_amdgcn_s_waitcnt(0xC07F);  // own reads of buffer kt % S complete
    lds_barrier();
    PSDP_DMA8_ISSUE((kt + S < KT ? kt + S : KT - 1) * BK, kt % S)
    const char* nxt = smem + ((kt + 1) % S) * SB;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      PSDP_KSTEP_MMA(ks)
#pragma unroll
      for (int p = 0; p < P; ++p) {
#pragma unroll
        for (int m = 0; m < MA; ++m) fa[ks][p][m] = frag_a(nxt, ks, p, m);
#pragma unroll
        for (int m = 0; m < MB; ++m) fb[ks][p][m] = frag_b(nxt, ks, p, m);
      }
    }
#pragma unroll
    for (int g = 0; g < NREAD; ++g) {
      if (g < L) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);            // DS read
      __builtin_amdgcn_sched_group_barrier(0x008, MR, 0);            // MFMA
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) { PSDP_KSTEP_MMA(ks) }
#undef PSDP_KSTEP_MMA
#undef PSDP_DMA8_ISSUE
  __builtin_amdgcn_sched_barrier(0);
  wait_vmcnt<0>();
  lds_barrier();
  float* const ptile = d.mode == EPI_PARTIAL ? d.part + ((int64_t)t.mat * d.ksplit + t.pad) * 65536 : nullptr;

  constexpr float PSCALE = (FMT == FMT_F16X3 || FMT == FMT_E4M3) ? F16X3_PROD_SCALE : 1.0f;
  // Two phases; in phase ph every wave stages its row blocks 2ph and 2ph+1,
  // i.e. tile rows wr*128 + ph*64 .. +64, at staged rows wr*64 .. +64, so that
  // only half of the accumulators stay live while an epilogue runs.
  float* Ct = (float*)smem;
  constexpr int CS = 256 + 1;
  // (phases written out: acc is indexed by compile-time constants only)
#define PSDP_STAGE8(PH)                                                                           \
  _Pragma("unroll") for (int h = 0; h < 2; ++h)                                                   \
  _Pragma("unroll") for (int mj = 0; mj < MB; ++mj)                                               \
  _Pragma("unroll") for (int e = 0; e < 16; ++e) {                                                \
    const int r = wr * 64 + h * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);                    \
    const int c = wc * 64 + mj * 32 + (lane & 31);                                                \
    Ct[r * CS + c] = acc[2 * (PH) + h][mj][e];                                                    \
  }                                                                                               \
  __syncthreads();                                                                                \
  if (ptile) {                                                                                    \
    partial_rows<256, NT>(ptile, Ct, (PH) * 64, 64);                                              \
    partial_rows<256, NT>(ptile, Ct + 64 * CS, 128 + (PH) * 64, 64);                              \
  } else {                                                                                        \
    epilogue_rows<256, NT, 2, true, FMT == FMT_F16X3>(d, bi, bj, Ct, (PH) * 64, 64, PSCALE);      \
    epilogue_rows<256, NT, 2, true, FMT == FMT_F16X3>(d, bi, bj, Ct + 64 * CS, 128 + (PH) * 64, 64, PSCALE); \
  }                                                                                               \
  __syncthreads();
  PSDP_STAGE8(0)
  PSDP_STAGE8(1)
#undef PSDP_STAGE8
}

template <int FMT>
__global__ __launch_bounds__(512) void sym_prod_dma8_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT, 8, false>::LDS_BYTES];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile_dma8<FMT>(d, t, smem);
}

// ------------------------------------------------------------------ FMT_F16MX
// Products of FMT_F16MX operands.  Per 32-k block and 32 x 32 accumulator:
// one scaled FP8 MFMA whose 64 k are the block's hi.lo (first 32) and lo.hi
// (last 32), then the block's two binary16 hi.hi MFMAs, all into one sum
// (smallest terms first, as FP16x3's lo.hi, hi.lo, hi.hi).  The operand
// pieces are staged like FP16x3's two pieces (same bytes per row and slab,
// same swizzles and fragment reads), plus each slab's block scales.
//
// v_mfma_scale_f32_32x32x64_f8f6f4 with e4m3 operands (measured on gfx950):
// lane l holds row l % 32; bytes 0-15 of lanes l and l + 32 are the 32 k of
// the first block, bytes 16-31 those of the second, A and B lanes pairing
// byte for byte; the scale operand (low byte) of lane r + 32 b scales block
// b of row r.  So A lane (r, h) passes hi bytes 16h.. then lo bytes 16h..
// with its row's hi (h = 0) or lo (h = 1) scale, and B lane (c, h) passes lo
// bytes then hi bytes with its row's lo (h = 0) or hi (h = 1) scale.
typedef int i32x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x16 mfma_mx(i32x8 a, i32x8 b, int sa, int sb, f32x16 c) {
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
}
// The 32 bytes of an e4m3 fragment from two 16-byte LDS reads, built where
// they are read so that they land in one register tuple.
__device__ __forceinline__ i32x8 cat16(u32x4 x, u32x4 y) {
  return __builtin_bit_cast(i32x8, __builtin_shufflevector(x, y, 0, 1, 2, 3, 4, 5, 6, 7));
}

// Scale words of a stage: thread t (mod the stage's word count) copies 4
// bytes (two rows) of segment (panel g, block q): rows of the X panel of bi
// for g = 0, of the Y panel of bj for g = 1.  With more threads than words
// the extra threads copy the same words again (same bytes to the same place),
// so that every wave issues the same number of loads.
template <int BT, int NQ, int NT>
struct MxScaleLoad {
  static constexpr int SCB = 2 * 2 * BT * NQ, SW = SCB / 4;
  static_assert(SW % 64 == 0, "scale words per stage must fill whole waves");
  const uint16_t* src;
  int dst;  // wave-uniform byte offset in the stage's scale area
  __device__ __forceinline__ MxScaleLoad(const ProdDesc& d, int row0, int col0, int tid, int lane) {
    const int sw = tid % SW;
    const int seg = (4 * sw) / (2 * BT), row = ((4 * sw) % (2 * BT)) / 2;
    const int g = seg / NQ, q = seg % NQ;
    src = (const uint16_t*)(g == 0 ? d.x[2] : d.y[2]) + (int64_t)q * d.ld + (g == 0 ? row0 : col0) + row;
    dst = __builtin_amdgcn_readfirstlane(4 * (sw - lane));
  }
  __device__ __forceinline__ void issue(int kt, char* area, int64_t ld) const {
    __builtin_amdgcn_global_load_lds((const void*)(src + (int64_t)kt * NQ * ld), (lds_ptr_t)(area + dst), 4, 0, 0);
  }
};

// 4 waves, 2 x 2 of (16 SUB) x (16 SUB), M x M accumulators each (SUB = 2, 4).
template <int SUB>
__device__ __forceinline__ void sym_tile_dma_mx(const ProdDesc& d, int bi, int bj, char* smem) {
  using DT = DmaTiling<FMT_F16MX, SUB, false>;
  constexpr int P = 2, BK = DT::BK, RB = DT::RB, S = DT::S, L = DT::L, BT = DT::BT;
  constexpr int TB = DT::TILE_BYTES, SB = DT::STAGE_BYTES, SST = DT::STAGE_STRIDE;
  constexpr int NQ = BK / 32, LT = L + 1;  // blocks per slab, loads per thread per stage
  static_assert((S - 1) * LT <= 63, "vmcnt range");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = (wave & 3) >> 1, wc = wave & 1;
  const int64_t ld = d.ld;
  const int row0 = bi * BT, col0 = bj * BT;

  constexpr int CPR = RB / 16, RPB = 1024 / RB;
  const uint16_t* gsrc[L];
  int lofs[L], loff[L];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int q = wave * L + i;
    const int gp = q / DT::BLOCKS, blk = q % DT::BLOCKS;
    const int g = gp / P, p = gp % P;
    const void* bx = p == 0 ? d.x[0] : d.x[1];
    const void* by = p == 0 ? d.y[0] : d.y[1];
    const void* base = g == 0 ? bx : by;
    gsrc[i] = (const uint16_t*)base + (int64_t)((g == 0 ? row0 : col0) + blk * RPB) * ld;
    lofs[i] = gp * TB + blk * 1024;
    const int r = blk * RPB + lane / CPR;
    loff[i] = (lane / CPR) * (int)ld + ((lane % CPR) ^ dma_swz<RB>(r)) * 8;
  }
  const MxScaleLoad<BT, NQ, 256> sload(d, row0, col0, tid, lane);
#define PSDP_MX_ISSUE(KT, BUF)                                                                       \
  {                                                                                                 \
    _Pragma("unroll") for (int i = 0; i < L; ++i)                                                   \
        __builtin_amdgcn_global_load_lds((const void*)(gsrc[i] + (KT) * BK + loff[i]),              \
                                         (lds_ptr_t)(smem + (BUF) * SST + lofs[i]), 16, 0, 0);      \
    sload.issue(KT, smem + (BUF) * SST + SB, ld);                                                   \
  }

  constexpr int M = SUB / 2, KS = BK / 16;
  f32x16 acc[M][M];
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.f;

  const int arow = (wr * DT::WT + (lane & 31)) * RB;
  const int brow = (wc * DT::WT + (lane & 31)) * RB;
  const int swzl = dma_swz<RB>(lane & 31);
  const int h = lane >> 5;
  const int KT = d.npad / BK;

  // binary16 hi fragments per k-step, e4m3 fragments per block
  u32x4 fa[KS][M], fb[KS][M], na[KS][M];
  i32x8 fa8[NQ][M], fb8[NQ][M], na8[NQ][M];
  int sa[NQ][M], sb[NQ][M], nsa[NQ][M];
  auto frag_a = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + p * TB + arow + m * 32 * RB + 16 * ((ks * 2 + h) ^ swzl));
  };
  auto frag_b = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + (P + p) * TB + brow + m * 32 * RB + 16 * ((ks * 2 + h) ^ swzl));
  };
  // A: hi bytes (chunk of k-step 2q) then lo bytes (k-step 2q + 1); B: lo then hi
  auto frag_a8 = [&](const char* stg, int q, int m) -> i32x8 {
    return cat16(frag_a(stg, 2 * q, 1, m), frag_a(stg, 2 * q + 1, 1, m));
  };
  auto frag_b8 = [&](const char* stg, int q, int m) -> i32x8 {
    return cat16(frag_b(stg, 2 * q + 1, 1, m), frag_b(stg, 2 * q, 1, m));
  };
  // scale bytes: A lane (r, h) takes byte h of its row's pair, B lane byte 1 - h
  auto scl_a = [&](const char* stg, int q, int m) -> int {
    return *(const unsigned char*)(stg + SB + (q * BT + wr * DT::WT + m * 32 + (lane & 31)) * 2 + h);
  };
  auto scl_b = [&](const char* stg, int q, int m) -> int {
    return *(const unsigned char*)(stg + SB + ((NQ + q) * BT + wc * DT::WT + m * 32 + (lane & 31)) * 2 + 1 - h);
  };

#pragma unroll
  for (int st = 0; st < S; ++st) PSDP_MX_ISSUE((st < KT ? st : KT - 1), st)
  wait_vmcnt<(S - 1) * LT>();
  lds_barrier();
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int m = 0; m < M; ++m) {
      fa[ks][m] = frag_a(smem, ks, 0, m);
      fb[ks][m] = frag_b(smem, ks, 0, m);
    }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int m = 0; m < M; ++m) {
      fa8[q][m] = frag_a8(smem, q, m);
      fb8[q][m] = frag_b8(smem, q, m);
      sa[q][m] = scl_a(smem, q, m);
      sb[q][m] = scl_b(smem, q, m);
    }
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)

#define PSDP_MX_COLUMN(q, mj)                                                                     \
  _Pragma("unroll") for (int mi = 0; mi < M; ++mi) {                                              \
    f32x16 c = mfma_mx(fa8[q][mi], fb8[q][mj], sa[q][mi], sb[q][mj], acc[mi][mj]);                \
    c = mfma32(fa[2 * (q)][mi], fb[2 * (q)][mj], c, FMT_F16);                                     \
    acc[mi][mj] = mfma32(fa[2 * (q) + 1][mi], fb[2 * (q) + 1][mj], c, FMT_F16);                   \
  }

  constexpr int NREAD = 2 * KS * P * M;
  constexpr int NMFMA = NQ * M * M * 3;
  constexpr int MR = NMFMA / NREAD > 0 ? NMFMA / NREAD : 1;
  for (int kt = 0; kt + 1 < KT; ++kt) {
    wait_vmcnt<(S - 2) * LT>();          // own DMA of slab kt + 1 landed
    __builtin_amdgcn_s_waitcnt(0xC07F);  // own reads of buffer kt % S complete
    lds_barrier();
    PSDP_MX_ISSUE((kt + S < KT ? kt + S : KT - 1), kt % S)
    const char* nxt = smem + ((kt + 1) % S) * SST;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int m = 0; m < M; ++m) na[ks][m] = frag_a(nxt, ks, 0, m);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int m = 0; m < M; ++m) {
        na8[q][m] = frag_a8(nxt, q, m);
        nsa[q][m] = scl_a(nxt, q, m);
      }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int mj = 0; mj < M; ++mj) {
        PSDP_MX_COLUMN(q, mj)
        fb[2 * q][mj] = frag_b(nxt, 2 * q, 0, mj);
        fb[2 * q + 1][mj] = frag_b(nxt, 2 * q + 1, 0, mj);
        fb8[q][mj] = frag_b8(nxt, q, mj);
        sb[q][mj] = scl_b(nxt, q, mj);
      }
#pragma unroll
    for (int g = 0; g < NREAD; ++g) {
      if (g < L) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);            // DS read
      __builtin_amdgcn_sched_group_barrier(0x008, MR, 0);            // MFMA
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int m = 0; m < M; ++m) fa[ks][m] = na[ks][m];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int m = 0; m < M; ++m) {
        fa8[q][m] = na8[q][m];
        sa[q][m] = nsa[q][m];
      }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int mj = 0; mj < M; ++mj) { PSDP_MX_COLUMN(q, mj) }
#undef PSDP_MX_COLUMN
#undef PSDP_MX_ISSUE
  __builtin_amdgcn_sched_barrier(0);
  wait_vmcnt<0>();
  lds_barrier();

  float* Ct = (float*)smem;
  constexpr int CS = BT + 1;
  constexpr int PHASES = BT / DT::HR;
#pragma unroll
  for (int ph = 0; ph < PHASES; ++ph) {
    if (PHASES == 1 || wr == ph) {
#pragma unroll
      for (int mi = 0; mi < M; ++mi)
#pragma unroll
        for (int mj = 0; mj < M; ++mj)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int r = wr * DT::WT + mi * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3) - ph * DT::HR;
            const int c = wc * DT::WT + mj * 32 + (lane & 31);
            Ct[r * CS + c] = acc[mi][mj][e];
          }
    }
    __syncthreads();
    epilogue_rows<BT, 256>(d, bi, bj, Ct, ph * DT::HR, DT::HR, F16X3_PROD_SCALE);
    __syncthreads();
  }
}

// 256 x 256 tile on 8 waves (as sym_tile_dma8): wave w owns rows wr*128 ..
// +128 and columns wc*64 .. +64, 4 x 2 accumulators.  Per slab (one 32-k
// block): the 8 scaled FP8 MFMAs first, while the next slab's e4m3 fragments
// and scales are read into the registers they free, then the two k-steps of
// binary16 MFMAs, each followed by the next slab's hi fragments of that step.
__device__ __forceinline__ void sym_tile_dma8_mx(const ProdDesc& d, const TileRef& t, char* smem) {
  const int bi = t.bi, bj = t.bj;
  using DT = DmaTiling<FMT_F16MX, 8, false>;
  constexpr int P = 2, BK = DT::BK, RB = DT::RB, S = DT::S;
  constexpr int TB = DT::TILE_BYTES, SB = DT::STAGE_BYTES, SST = DT::STAGE_STRIDE;
  constexpr int NT = 512;
  constexpr int L = SB / (NT * 16);
  constexpr int LT = L + 1;
  static_assert(BK == 32 && SB % (NT * 16) == 0 && L >= 1, "one 32-k block per slab over 8 waves");
  static_assert((S - 1) * LT <= 63, "vmcnt range");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 2, wc = wave & 3;
  const int64_t ld = d.ld;
  const int row0 = bi * 256, col0 = bj * 256;

  constexpr int CPR = RB / 16, RPB = 1024 / RB;
  const uint16_t* gsrc[L];
  int lofs[L], loff[L];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int q = wave * L + i;
    const int gp = q / DT::BLOCKS, blk = q % DT::BLOCKS;
    const int g = gp / P, p = gp % P;
    const void* bx = p == 0 ? d.x[0] : d.x[1];
    const void* by = p == 0 ? d.y[0] : d.y[1];
    const void* base = g == 0 ? bx : by;
    gsrc[i] = (const uint16_t*)base + (int64_t)((g == 0 ? row0 : col0) + blk * RPB) * ld;
    lofs[i] = gp * TB + blk * 1024;
    const int r = blk * RPB + lane / CPR;
    loff[i] = (lane / CPR) * (int)ld + ((lane % CPR) ^ dma_swz<RB>(r)) * 8;
  }
  MxScaleLoad<256, 1, NT> sload(d, row0, col0, tid, lane);
#define PSDP_MX8_ISSUE(KT, BUF)                                                                      \
  {                                                                                                 \
    _Pragma("unroll") for (int i = 0; i < L; ++i)                                                   \
        __builtin_amdgcn_global_load_lds((const void*)(gsrc[i] + (KT) * BK + loff[i]),              \
                                         (lds_ptr_t)(smem + (BUF) * SST + lofs[i]), 16, 0, 0);      \
    sload.issue(KT, smem + (BUF) * SST + SB, ld);                                                   \
  }

  constexpr int MA = 4, MB = 2;
  f32x16 acc[MA][MB];
#pragma unroll
  for (int a = 0; a < MA; ++a)
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.f;

  const int arow = (wr * 128 + (lane & 31)) * RB;
  const int brow = (wc * 64 + (lane & 31)) * RB;
  const int swzl = dma_swz<RB>(lane & 31);
  const int h = lane >> 5;
  int k0;
  const int KT = tile_k_range<BK>(d, t.pad, k0);
#pragma unroll
  for (int i = 0; i < L; ++i) gsrc[i] += (int64_t)k0 * BK;
  sload.src += (int64_t)k0 * ld;  // one scale row per 32-k block

  u32x4 fa[2][MA], fb[2][MB];  // binary16 hi fragments per k-step
  i32x8 fa8[MA], fb8[MB];      // e4m3 fragments of the slab's block
  int sa[MA], sb[MB];
  auto frag_a = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + p * TB + arow + m * 32 * RB + 16 * ((ks * 2 + h) ^ swzl));
  };
  auto frag_b = [&](const char* stg, int ks, int p, int m) -> u32x4 {
    return *(const u32x4*)(stg + (P + p) * TB + brow + m * 32 * RB + 16 * ((ks * 2 + h) ^ swzl));
  };
  auto frag_a8 = [&](const char* stg, int m) -> i32x8 { return cat16(frag_a(stg, 0, 1, m), frag_a(stg, 1, 1, m)); };
  auto frag_b8 = [&](const char* stg, int m) -> i32x8 { return cat16(frag_b(stg, 1, 1, m), frag_b(stg, 0, 1, m)); };
  auto scl_a = [&](const char* stg, int m) -> int {
    return *(const unsigned char*)(stg + SB + (wr * 128 + m * 32 + (lane & 31)) * 2 + h);
  };
  auto scl_b = [&](const char* stg, int m) -> int {
    return *(const unsigned char*)(stg + SB + (256 + wc * 64 + m * 32 + (lane & 31)) * 2 + 1 - h);
  };

#pragma unroll
  for (int st = 0; st < S; ++st) PSDP_MX8_ISSUE((st < KT ? st : KT - 1), st)
  wait_vmcnt<(S - 1) * LT>();
  lds_barrier();
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
    for (int m = 0; m < MA; ++m) fa[ks][m] = frag_a(smem, ks, 0, m);
#pragma unroll
    for (int m = 0; m < MB; ++m) fb[ks][m] = frag_b(smem, ks, 0, m);
  }
#pragma unroll
  for (int m = 0; m < MA; ++m) fa8[m] = frag_a8(smem, m), sa[m] = scl_a(smem, m);
#pragma unroll
  for (int m = 0; m < MB; ++m) fb8[m] = frag_b8(smem, m), sb[m] = scl_b(smem, m);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)

#define PSDP_MX8_CROSS()                                                                          \
  _Pragma("unroll") for (int mj = 0; mj < MB; ++mj)                                               \
  _Pragma("unroll") for (int mi = 0; mi < MA; ++mi)                                               \
    acc[mi][mj] = mfma_mx(fa8[mi], fb8[mj], sa[mi], sb[mj], acc[mi][mj]);
#define PSDP_MX8_HH(ks)                                                                           \
  _Pragma("unroll") for (int mj = 0; mj < MB; ++mj)                                               \
  _Pragma("unroll") for (int mi = 0; mi < MA; ++mi)                                               \
    acc[mi][mj] = mfma32(fa[ks][mi], fb[ks][mj], acc[mi][mj], FMT_F16);

  // reads per slab: 24 fragments + 6 scale bytes; MFMAs: 8 FP8 + 16 binary16
  for (int kt = 0; kt + 1 < KT; ++kt) {
    wait_vmcnt<(S - 2) * LT>();          // own DMA of slab kt + 1 landed
    __builtin_amdgcn_s_waitcnt(0xC07F);  // own reads of buffer kt % S complete
    lds_barrier();
    PSDP_MX8_ISSUE((kt + S < KT ? kt + S : KT - 1), kt % S)
    const char* nxt = smem + ((kt + 1) % S) * SST;
    PSDP_MX8_CROSS()
#pragma unroll
    for (int m = 0; m < MA; ++m) fa8[m] = frag_a8(nxt, m), sa[m] = scl_a(nxt, m);
#pragma unroll
    for (int m = 0; m < MB; ++m) fb8[m] = frag_b8(nxt, m), sb[m] = scl_b(nxt, m);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      PSDP_MX8_HH(ks)
#pragma unroll
      for (int m = 0; m < MA; ++m) fa[ks][m] = frag_a(nxt, ks, 0, m);
#pragma unroll
      for (int m = 0; m < MB; ++m) fb[ks][m] = frag_b(nxt, ks, 0, m);
    }
#pragma unroll
    for (int g = 0; g < 24; ++g) {
      if (g < LT) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);             // DS read
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);             // MFMA
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  PSDP_MX8_CROSS()
  PSDP_MX8_HH(0)
  PSDP_MX8_HH(1)
#undef PSDP_MX8_CROSS
#undef PSDP_MX8_HH
#undef PSDP_MX8_ISSUE
  __builtin_amdgcn_sched_barrier(0);
  wait_vmcnt<0>();
  lds_barrier();
  float* const ptile = d.mode == EPI_PARTIAL ? d.part + ((int64_t)t.mat * d.ksplit + t.pad) * 65536 : nullptr;

  float* Ct = (float*)smem;
  constexpr int CS = 256 + 1;
#define PSDP_STAGE8(PH)                                                                           \
  _Pragma("unroll") for (int hh = 0; hh < 2; ++hh)                                                \
  _Pragma("unroll") for (int mj = 0; mj < MB; ++mj)                                               \
  _Pragma("unroll") for (int e = 0; e < 16; ++e) {                                                \
    const int r = wr * 64 + hh * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);                   \
    const int c = wc * 64 + mj * 32 + (lane & 31);                                                \
    Ct[r * CS + c] = acc[2 * (PH) + hh][mj][e];                                                   \
  }                                                                                               \
  __syncthreads();                                                                                \
  if (ptile) {                                                                                    \
    partial_rows<256, NT>(ptile, Ct, (PH) * 64, 64);                                              \
    partial_rows<256, NT>(ptile, Ct + 64 * CS, 128 + (PH) * 64, 64);                              \
  } else {                                                                                        \
    epilogue_rows<256, NT, 2>(d, bi, bj, Ct, (PH) * 64, 64, F16X3_PROD_SCALE);                    \
    epilogue_rows<256, NT, 2>(d, bi, bj, Ct + 64 * CS, 128 + (PH) * 64, 64, F16X3_PROD_SCALE);    \
  }                                                                                               \
  __syncthreads();
  PSDP_STAGE8(0)
  PSDP_STAGE8(1)
#undef PSDP_STAGE8
}

__global__ __launch_bounds__(512) void sym_prod_dma8_mx_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT_F16MX, 8, false>::LDS_BYTES];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile_dma8_mx(d, t, smem);
}

// Second half of a split product: SPLIT_REDUCE_WGS workgroups per split
// 256 x 256 tile ({slot, bi, bj, -}), one per SPLIT_REDUCE_ROWS rows, add the
// tile's d.ksplit partial tiles in chunk order (((p0 + p1) + p2) + ...), scale
// the sum by cscale (the 2^-16 of FP16x3 / F16MX products, inside the
// epilogue) and run the product's epilogue (mode d.mode) on their rows.  A
// thread sums 4 quads at a time and keeps 4 chunks of loads in flight
// (clamped addresses past the last chunk, whose values are not added), so the
// reads do not wait one by one.  32 rows per workgroup keep every F16MX block
// (32 values of a row, also of the mirrored rows) inside one workgroup.
constexpr int SPLIT_REDUCE_ROWS = 32, SPLIT_REDUCE_WGS = 256 / SPLIT_REDUCE_ROWS;
template <bool A3PCS>
__global__ __launch_bounds__(256) void sym_split_reduce_kernel(ProdDesc d, const TileRef* __restrict__ tiles,
                                                               float cscale) {
  __shared__ __attribute__((aligned(16))) float Ct[SPLIT_REDUCE_ROWS * 257];
  const TileRef t = tiles[blockIdx.x / SPLIT_REDUCE_WGS];
  const int part = blockIdx.x % SPLIT_REDUCE_WGS;
  if (t.bi < 0) return;
  const int S = d.ksplit;
  auto src = gbl((const float*)d.part) + (int64_t)t.mat * S * 65536 + part * SPLIT_REDUCE_ROWS * 256;
  constexpr int G = 4, KB = 4;
  static_assert(SPLIT_REDUCE_ROWS % 32 == 0 && (SPLIT_REDUCE_ROWS * 64) % (256 * G) == 0,
                "whole F16MX blocks per workgroup; every quad index of the loop in range");
  typedef __attribute__((address_space(1))) const f32x4 gf4;
  for (int q0 = threadIdx.x; q0 < SPLIT_REDUCE_ROWS * 64; q0 += 256 * G) {
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g * 256;
      acc[g] = *(gf4*)(src + (q >> 6) * 256 + (q & 63) * 4);
    }
    for (int k0 = 1; k0 < S; k0 += KB) {
      f32x4 w[KB][G];
#pragma unroll
      for (int kk = 0; kk < KB; ++kk) {
        const int k = min(k0 + kk, S - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int q = q0 + g * 256;
          w[kk][g] = *(gf4*)(src + (int64_t)k * 65536 + (q >> 6) * 256 + (q & 63) * 4);
        }
      }
#pragma unroll
      for (int kk = 0; kk < KB; ++kk) {
        if (k0 + kk < S) {
#pragma unroll
          for (int g = 0; g < G; ++g) {
            acc[g].x += w[kk][g].x;
            acc[g].y += w[kk][g].y;
            acc[g].z += w[kk][g].z;
            acc[g].w += w[kk][g].w;
          }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g * 256;
      float* row = Ct + (q >> 6) * 257 + (q & 63) * 4;
      row[0] = acc[g].x;
      row[1] = acc[g].y;
      row[2] = acc[g].z;
      row[3] = acc[g].w;
    }
  }
  __syncthreads();
  epilogue_rows<256, 256, 8, true, A3PCS>(d, t.bi, t.bj, Ct, part * SPLIT_REDUCE_ROWS, SPLIT_REDUCE_ROWS, cscale);
}

template <int SUB>
__global__ __launch_bounds__(256) void sym_prod_dma_mx_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT_F16MX, SUB, false>::LDS_BYTES];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile_dma_mx<SUB>(d, t.bi, t.bj, smem);
}

template <int SUB>
__global__ __launch_bounds__(256) void sym_prod_dma_mx_grouped_kernel(const ProdDesc* __restrict__ descs,
                                                                      const TileRef* __restrict__ tiles,
                                                                      int ntiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT_F16MX, SUB, false>::LDS_BYTES];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    if (tr.mat < 0) continue;
    sym_tile_dma_mx<SUB>(descs[tr.mat], tr.bi, tr.bj, smem);
  }
}

template <int FMT, int SUB>
__global__ __launch_bounds__(256) void sym_prod_dma_grouped_kernel(const ProdDesc* __restrict__ descs,
                                                                   const TileRef* __restrict__ tiles, int ntiles) {
  __shared__ __attribute__((aligned(16))) char smem[DmaTiling<FMT, SUB>::LDS_BYTES];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    if (tr.mat < 0) continue;  // padding entry of an XCD-interleaved list
    sym_tile_dma<FMT, SUB, false>(descs[tr.mat], tr.bi, tr.bj, smem);  // descriptor read in place
  }
}

// One workgroup per upper-triangle tile of a single matrix.
template <int FMT, int SUB, bool AVG>
__global__ __launch_bounds__(256) void sym_prod_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  using TL = Tiling<FMT, SUB, AVG>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[TL::SMEM_ELEMS];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile<FMT, SUB, AVG>(d, t.bi, t.bj, smem);
}

// Persistent grouped launch: workgroups stride over a tile list that spans
// many matrices (tiles sorted by cost on the host).
template <int FMT, int SUB, bool AVG>
__global__ __launch_bounds__(256) void sym_prod_grouped_kernel(const ProdDesc* __restrict__ descs,
                                                               const TileRef* __restrict__ tiles,
                                                               int ntiles) {
  using TL = Tiling<FMT, SUB, AVG>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[TL::SMEM_ELEMS];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    if (tr.mat < 0) continue;  // padding entry of an XCD-interleaved list
    const ProdDesc d = descs[tr.mat];
    sym_tile<FMT, SUB, AVG>(d, tr.bi, tr.bj, smem);
  }
}

// Float32 tiles keep two workgroups per CU (8 waves): their two-level sums
// (total + partial accumulators) must then fit 256 registers per wave, which
// the compiler is held to here (without the bound it took 274 and ran one
// workgroup per CU, 19% slower at n = 16384).
template <int SUB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void sym_prod_f32_kernel(
    ProdDesc d, const TileRef* __restrict__ tiles) {
  using TL = Tiling<FMT_F32, SUB, false>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[TL::SMEM_ELEMS];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile<FMT_F32, SUB, false>(d, t.bi, t.bj, smem);
}

template <int SUB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void sym_prod_f32_grouped_kernel(
    const ProdDesc* __restrict__ descs, const TileRef* __restrict__ tiles, int ntiles) {
  using TL = Tiling<FMT_F32, SUB, false>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[TL::SMEM_ELEMS];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    if (tr.mat < 0) continue;
    const ProdDesc d = descs[tr.mat];
    sym_tile<FMT_F32, SUB, false>(d, tr.bi, tr.bj, smem);
  }
}

template <int FMT, int SUB, bool AVG>
static hipError_t launch_one(const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s) {
  if constexpr (FMT == FMT_F32 && !AVG)
    hipLaunchKernelGGL((sym_prod_f32_kernel<SUB>), dim3(ntiles), dim3(256), 0, s, d, tiles);
  else
    hipLaunchKernelGGL((sym_prod_kernel<FMT, SUB, AVG>), dim3(ntiles), dim3(256), 0, s, d, tiles);
  return hipGetLastError();
}

template <int FMT, int SUB, bool AVG>
static hipError_t launch_grouped(const ProdDesc* descs, const TileRef* tiles, int ntiles, int grid,
                                 hipStream_t s) {
  if constexpr (FMT == FMT_F32 && !AVG)
    hipLaunchKernelGGL((sym_prod_f32_grouped_kernel<SUB>), dim3(grid), dim3(256), 0, s, descs, tiles, ntiles);
  else
    hipLaunchKernelGGL((sym_prod_grouped_kernel<FMT, SUB, AVG>), dim3(grid), dim3(256), 0, s, descs, tiles,
                       ntiles);
  return hipGetLastError();
}

// 256 x 256 tiles (SUB = 8) exist for the 1- and 2-piece 16-bit formats with
// plain (non-averaged) products.
bool sym_big_tile_capable(int fmt, int avg) {
  return !avg && (fmt == FMT_F16 || fmt == FMT_BF16 || fmt == FMT_F16X3 || fmt == FMT_BF16X3 || fmt == FMT_E4M3 ||
                  fmt == FMT_F16MX);
}

// Producer waves (WS) are kept available but off: measured at n = 16384 they
// ran the 128 x 128 FP16x3 tiles 5% slower than the plain kernel, whose
// 64-element slabs do not fit the 256 registers a second wave per SIMD leaves.
// 256 x 256 tiles of the two-piece formats run on 8 waves (sym_tile_dma8):
// measured at n = 16384 (alternating runs on one device) a binary16x3
// product takes 11.31-11.40 ms against 12.10-12.15 ms with 4 waves.  One-piece
// formats keep 4 waves: with half the MFMAs per slab the 8-wave layout's
// fragment reads (24 KB per wave per slab) make LDS the limit (binary16:
// 4.31 ms against 4.23 ms).
template <int FMT, int SUB>
static hipError_t launch_dma(const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s) {
  if constexpr (SUB == 8 && FmtTraits<FMT>::P == 2) {
    hipLaunchKernelGGL((sym_prod_dma8_kernel<FMT>), dim3(ntiles), dim3(512), 0, s, d, tiles);
    return hipGetLastError();
  }
  constexpr bool WS = false;
  if constexpr (SUB == 8) {
    if (d.mode == EPI_PARTIAL) {
      hipLaunchKernelGGL((sym_prod_dma_kernel<FMT, 8, WS, true>), dim3(ntiles), dim3(256), 0, s, d, tiles);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((sym_prod_dma_kernel<FMT, SUB, WS>), dim3(ntiles), dim3(WS ? 512 : 256), 0, s, d, tiles);
  return hipGetLastError();
}

template <int FMT, int SUB>
static hipError_t launch_dma_grouped(const ProdDesc* descs, const TileRef* tiles, int ntiles, int grid,
                                     hipStream_t s) {
  hipLaunchKernelGGL((sym_prod_dma_grouped_kernel<FMT, SUB>), dim3(grid), dim3(256), 0, s, descs, tiles, ntiles);
  return hipGetLastError();
}

// avg selects the averaged (Jordan) product; it exists for the 1-piece
// 16-bit formats, whose operand rounding makes C visibly non-symmetric.
// Plain 16-bit products use the direct-to-LDS kernels; FP32 and averaged
// products use the register-staged ones.
hipError_t launch_sym_product(int fmt, int sub, int avg, const ProdDesc& d, const TileRef* tiles, int ntiles,
                              hipStream_t s) {
  if (avg) {
    if (fmt == FMT_F16)
      return sub == 4 ? launch_one<FMT_F16, 4, true>(d, tiles, ntiles, s)
                      : (sub == 2 ? launch_one<FMT_F16, 2, true>(d, tiles, ntiles, s) : hipErrorInvalidValue);
    if (fmt == FMT_BF16)
      return sub == 4 ? launch_one<FMT_BF16, 4, true>(d, tiles, ntiles, s)
                      : (sub == 2 ? launch_one<FMT_BF16, 2, true>(d, tiles, ntiles, s) : hipErrorInvalidValue);
    return hipErrorInvalidValue;
  }
#define PSDP_CASE1(F)                                                           \
  case F:                                                                       \
    return sub == 8 ? launch_dma<F, 8>(d, tiles, ntiles, s)                     \
                    : (sub == 4 ? launch_dma<F, 4>(d, tiles, ntiles, s)         \
                                : launch_dma<F, 2>(d, tiles, ntiles, s));
#define PSDP_CASEN(F)                                                           \
  case F:                                                                       \
    return sub == 4 ? launch_dma<F, 4>(d, tiles, ntiles, s)                     \
                    : (sub == 2 ? launch_dma<F, 2>(d, tiles, ntiles, s) : hipErrorInvalidValue);
#define PSDP_CASE2(F)                                                           \
  case F:                                                                       \
    return sub == 8 ? launch_dma<F, 8>(d, tiles, ntiles, s)                     \
                    : (sub == 4 ? launch_dma<F, 4>(d, tiles, ntiles, s)         \
                                : (sub == 2 ? launch_dma<F, 2>(d, tiles, ntiles, s) : hipErrorInvalidValue));
  switch (fmt) {
    case FMT_F32:
      return sub == 4 ? launch_one<FMT_F32, 4, false>(d, tiles, ntiles, s)
                      : (sub == 2 ? launch_one<FMT_F32, 2, false>(d, tiles, ntiles, s) : hipErrorInvalidValue);
    PSDP_CASE1(FMT_F16)
    PSDP_CASE1(FMT_BF16)
    PSDP_CASE1(FMT_E4M3)
    PSDP_CASE2(FMT_F16X3)
    PSDP_CASE2(FMT_BF16X3)
    PSDP_CASEN(FMT_BF16X9)
    case FMT_F16MX:
      if (sub == 8) {
        hipLaunchKernelGGL(sym_prod_dma8_mx_kernel, dim3(ntiles), dim3(512), 0, s, d, tiles);
      } else if (sub == 4) {
        hipLaunchKernelGGL((sym_prod_dma_mx_kernel<4>), dim3(ntiles), dim3(256), 0, s, d, tiles);
      } else if (sub == 2) {
        hipLaunchKernelGGL((sym_prod_dma_mx_kernel<2>), dim3(ntiles), dim3(256), 0, s, d, tiles);
      } else {
        return hipErrorInvalidValue;
      }
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
#undef PSDP_CASE1
#undef PSDP_CASEN
#undef PSDP_CASE2
}

hipError_t launch_split_reduce(int fmt, const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s) {
  const float cscale = (fmt == FMT_F16X3 || fmt == FMT_E4M3 || fmt == FMT_F16MX) ? F16X3_PROD_SCALE : 1.0f;
  if (fmt == FMT_F16X3)
    hipLaunchKernelGGL(sym_split_reduce_kernel<true>, dim3(SPLIT_REDUCE_WGS * ntiles), dim3(256), 0, s, d, tiles, cscale);
  else
    hipLaunchKernelGGL(sym_split_reduce_kernel<false>, dim3(SPLIT_REDUCE_WGS * ntiles), dim3(256), 0, s, d, tiles, cscale);
  return hipGetLastError();
}

hipError_t launch_sym_product_grouped(int fmt, int sub, int avg, const ProdDesc* descs, const TileRef* tiles,
                                      int ntiles, int grid, hipStream_t s) {
  if (sub == 4 && !avg) {
    switch (fmt) {
      case FMT_F16: return launch_dma_grouped<FMT_F16, 4>(descs, tiles, ntiles, grid, s);
      case FMT_F16X3: return launch_dma_grouped<FMT_F16X3, 4>(descs, tiles, ntiles, grid, s);
      case FMT_BF16: return launch_dma_grouped<FMT_BF16, 4>(descs, tiles, ntiles, grid, s);
      case FMT_BF16X3: return launch_dma_grouped<FMT_BF16X3, 4>(descs, tiles, ntiles, grid, s);
      case FMT_BF16X9: return launch_dma_grouped<FMT_BF16X9, 4>(descs, tiles, ntiles, grid, s);
      case FMT_E4M3: return launch_dma_grouped<FMT_E4M3, 4>(descs, tiles, ntiles, grid, s);
      case FMT_F16MX:
        hipLaunchKernelGGL((sym_prod_dma_mx_grouped_kernel<4>), dim3(grid), dim3(256), 0, s, descs, tiles, ntiles);
        return hipGetLastError();
      default: break;  // FP32 and averaged products: the 64 x 64 kernels below
    }
  }
  if (sub == 4) {
    if (avg) {
      if (fmt == FMT_F16) return launch_grouped<FMT_F16, 4, true>(descs, tiles, ntiles, grid, s);
      if (fmt == FMT_BF16) return launch_grouped<FMT_BF16, 4, true>(descs, tiles, ntiles, grid, s);
      return hipErrorInvalidValue;
    }
    if (fmt == FMT_F32) return launch_grouped<FMT_F32, 4, false>(descs, tiles, ntiles, grid, s);
    return hipErrorInvalidValue;
  }
  if (sub != 2) return hipErrorInvalidValue;
  if (avg) {
    if (fmt == FMT_F16) return launch_grouped<FMT_F16, 2, true>(descs, tiles, ntiles, grid, s);
    if (fmt == FMT_BF16) return launch_grouped<FMT_BF16, 2, true>(descs, tiles, ntiles, grid, s);
    return hipErrorInvalidValue;
  }
  switch (fmt) {
    case FMT_F32: return launch_grouped<FMT_F32, 2, false>(descs, tiles, ntiles, grid, s);
    case FMT_F16: return launch_dma_grouped<FMT_F16, 2>(descs, tiles, ntiles, grid, s);
    case FMT_F16X3: return launch_dma_grouped<FMT_F16X3, 2>(descs, tiles, ntiles, grid, s);
    case FMT_BF16: return launch_dma_grouped<FMT_BF16, 2>(descs, tiles, ntiles, grid, s);
    case FMT_BF16X3: return launch_dma_grouped<FMT_BF16X3, 2>(descs, tiles, ntiles, grid, s);
    case FMT_BF16X9: return launch_dma_grouped<FMT_BF16X9, 2>(descs, tiles, ntiles, grid, s);
    case FMT_E4M3: return launch_dma_grouped<FMT_E4M3, 2>(descs, tiles, ntiles, grid, s);
    case FMT_F16MX:
      hipLaunchKernelGGL((sym_prod_dma_mx_grouped_kernel<2>), dim3(grid), dim3(256), 0, s, descs, tiles, ntiles);
      return hipGetLastError();
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psdp
