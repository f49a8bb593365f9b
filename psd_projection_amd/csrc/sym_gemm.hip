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
// replaces the reference's explicit symmetrize passes).
//
// Tiling: 256 threads = 2 x 2 waves; each wave owns a (16*SUB) x (16*SUB)
// sub-block made of 16x16 MFMA tiles.  K is staged through LDS in BK-wide
// slabs, double buffered (one barrier per slab).  Rows are padded by 16 bytes
// so that the 16-byte fragment reads of 16 consecutive rows hit distinct banks.
#include "psdp_device.h"

namespace psdp {

template <int FMT>
struct FmtTraits;
template <>
struct FmtTraits<FMT_F32> {
  using E = float;
  static constexpr int P = 1, NACC = 1, BK = 32;
};
template <>
struct FmtTraits<FMT_F16> {
  using E = uint16_t;
  static constexpr int P = 1, NACC = 1, BK = 64;
};
template <>
struct FmtTraits<FMT_F16X3> {
  using E = uint16_t;
  static constexpr int P = 2, NACC = 2, BK = 32;
};
template <>
struct FmtTraits<FMT_BF16> {
  using E = uint16_t;
  static constexpr int P = 1, NACC = 1, BK = 64;
};
template <>
struct FmtTraits<FMT_BF16X3> {
  using E = uint16_t;
  static constexpr int P = 2, NACC = 1, BK = 32;
};
template <>
struct FmtTraits<FMT_BF16X9> {
  using E = uint16_t;
  static constexpr int P = 3, NACC = 1, BK = 32;
};

template <int FMT, int SUB, bool AVG = false>
struct Tiling {
  using F = FmtTraits<FMT>;
  static constexpr int NPROD = AVG ? 2 : 1;    // averaged products keep C and C^T
  using E = typename F::E;
  static constexpr int ESZ = (int)sizeof(E);
  static constexpr int WT = 16 * SUB;          // wave tile
  static constexpr int BT = 2 * WT;            // block tile
  static constexpr int BK = F::BK;
  static constexpr int P = F::P;
  static constexpr int NACC = F::NACC;
  static constexpr int VEC = 16 / ESZ;         // elements per 16-byte chunk
  static constexpr int ROW = BK + VEC;         // padded LDS row (elements)
  static constexpr int TILE = BT * ROW;        // elements of one piece tile
  static constexpr int CPR = BK / VEC;         // chunks per row
  static constexpr int CHUNKS = BT * CPR;
  static constexpr int LOADS = CHUNKS / 256;   // chunks per thread per piece tile
  static constexpr int STAGE = 2 * NPROD * P * TILE;  // X_I, Y_J (then Y_I, X_J)
  static constexpr int LDS_MAIN = 2 * STAGE * ESZ;
  static constexpr int LDS_CTILE = BT * (BT + 1) * 4;
  static constexpr int LDS_BYTES = LDS_MAIN > LDS_CTILE ? LDS_MAIN : LDS_CTILE;
  static constexpr int SMEM_ELEMS = (LDS_BYTES + ESZ - 1) / ESZ;
  static_assert(CHUNKS % 256 == 0, "tile too small for 256 threads");
};

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c, int fmt) {
  if (fmt == FMT_BF16 || fmt == FMT_BF16X3 || fmt == FMT_BF16X9)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ void store_pieces4(int fmt, uint16_t* const* outp, int64_t idx, const float w[4]) {
  uint16_t p[4][3];
#pragma unroll
  for (int e = 0; e < 4; ++e) split_value(fmt, w[e], p[e]);
  const int np = fmt_pieces(fmt);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q < np) {
      uint2 pk;
      pk.x = (uint32_t)p[0][q] | ((uint32_t)p[1][q] << 16);
      pk.y = (uint32_t)p[2][q] | ((uint32_t)p[3][q] << 16);
      *(uint2*)(outp[q] + idx) = pk;
    }
  }
}

__device__ __forceinline__ void store_piece1(int fmt, uint16_t* const* outp, int64_t idx, float w) {
  uint16_t p[3];
  split_value(fmt, w, p);
  const int np = fmt_pieces(fmt);
  outp[0][idx] = p[0];
  if (np > 1) outp[1][idx] = p[1];
  if (np > 2) outp[2][idx] = p[2];
}

// Epilogue from a BT x BT float tile C staged in LDS (row stride BT + 1).
// Two passes: the primary block (i, j) is written along j, the mirror block
// (j, i) along i, so both are coalesced 16-byte stores.  Diagonal tiles write
// only i <= j in the first pass and i < j in the second.
template <int BT>
__device__ __forceinline__ void epilogue_from_lds(const ProdDesc& d, int bi, int bj, float* C) {
  constexpr int CS = BT + 1;
  constexpr int Q4 = BT / 4;
  const int tid = threadIdx.x;
  const int row0 = bi * BT, col0 = bj * BT;
  const bool diag = bi == bj;
  const int64_t ld = d.ld;
  if (d.mode == EPI_OUTPUT) {
    const double s = d.scale ? *d.scale : 1.0;
    const int64_t old = d.out_ld;
    // (i, j) at i + j*ld: contiguous along i
    for (int q = tid; q < BT * Q4; q += 256) {
      const int c = q / Q4, r0 = (q % Q4) * 4;
      const int j = col0 + c;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = row0 + r0 + e;
        if (i < d.n && j < d.n && (!diag || i <= j)) {
          const double v = (double)C[(r0 + e) * CS + c] * s;
          if (d.out_f64) ((double*)d.out)[(int64_t)i + (int64_t)j * old] = v;
          else ((float*)d.out)[(int64_t)i + (int64_t)j * old] = (float)v;
        }
      }
    }
    // (j, i) at j + i*ld: contiguous along j
    for (int q = tid; q < BT * Q4; q += 256) {
      const int r = q / Q4, c0 = (q % Q4) * 4;
      const int i = row0 + r;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = col0 + c0 + e;
        if (i < d.n && j < d.n && (!diag || i < j)) {
          const double v = (double)C[r * CS + c0 + e] * s;
          if (d.out_f64) ((double*)d.out)[(int64_t)j + (int64_t)i * old] = v;
          else ((float*)d.out)[(int64_t)j + (int64_t)i * old] = (float)v;
        }
      }
    }
    return;
  }
  const bool poly = d.mode == EPI_POLY || d.mode == EPI_POLY_HALF_ID;
  const bool half_id = d.mode == EPI_POLY_HALF_ID;
  const bool pieces = d.out_fmt != FMT_F32;
  // pass 1: primary block, rows of the tile
  for (int q = tid; q < BT * Q4; q += 256) {
    const int r = q / Q4, c0 = (q % Q4) * 4;
    const int i = row0 + r, j0 = col0 + c0;
    const int64_t idx = (int64_t)i * ld + j0;
    float w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = C[r * CS + c0 + e];
    if (poly) {
      const f32x4 a4 = *(const f32x4*)(d.c32 + idx);
      const f32x4 p4 = *(const f32x4*)(d.p3 + idx);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float x = d.a * a4[e];
        x = fmaf(d.b, p4[e], x);
        x = fmaf(d.c, w[e], x);
        if (half_id) x = (x + (i == j0 + e ? 1.0f : 0.0f)) * 0.5f;
        w[e] = x;
      }
    }
    if (!diag) {
      if (d.c32) *(f32x4*)(d.c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
      if (pieces) store_pieces4(d.out_fmt, d.outp, idx, w);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (i <= j0 + e) {
          if (d.c32) d.c32[idx + e] = w[e];
          if (pieces) store_piece1(d.out_fmt, d.outp, idx + e, w[e]);
        }
      }
    }
    if (poly) {
#pragma unroll
      for (int e = 0; e < 4; ++e) C[r * CS + c0 + e] = w[e];
    }
  }
  __syncthreads();
  // pass 2: mirror block, columns of the tile
  for (int q = tid; q < BT * Q4; q += 256) {
    const int c = q / Q4, r0 = (q % Q4) * 4;
    const int j = col0 + c, i0 = row0 + r0;
    const int64_t idx = (int64_t)j * ld + i0;
    float w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = C[(r0 + e) * CS + c];
    if (!diag) {
      if (d.c32) *(f32x4*)(d.c32 + idx) = f32x4{w[0], w[1], w[2], w[3]};
      if (pieces) store_pieces4(d.out_fmt, d.outp, idx, w);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (i0 + e < j) {
          if (d.c32) d.c32[idx + e] = w[e];
          if (pieces) store_piece1(d.out_fmt, d.outp, idx + e, w[e]);
        }
      }
    }
  }
}

// MFMAs over one BK-wide LDS slab (X pieces then Y pieces).
template <int FMT, int SUB>
__device__ __forceinline__ void mma_slab(const typename Tiling<FMT, SUB>::E* X,
                                         const typename Tiling<FMT, SUB>::E* Y, int arow, int brow, int lane,
                                         f32x4 (&acc)[Tiling<FMT, SUB>::NACC][SUB][SUB]) {
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
            acc[0][si][sj] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[si][q], bf[sj][q], acc[0][si][sj], 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      u32x4 af[P][SUB], bf[P][SUB];
      const int ko = kk + (lane >> 4) * 8;
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int s = 0; s < SUB; ++s) {
          af[p][s] = *(const u32x4*)(X + p * TILE + arow + s * 16 * ROW + ko);
          bf[p][s] = *(const u32x4*)(Y + p * TILE + brow + s * 16 * ROW + ko);
        }
#pragma unroll
      for (int si = 0; si < SUB; ++si)
#pragma unroll
        for (int sj = 0; sj < SUB; ++sj) {
          if constexpr (FMT == FMT_F16 || FMT == FMT_BF16) {
            acc[0][si][sj] = mfma16(af[0][si], bf[0][sj], acc[0][si][sj], FMT);
          } else if constexpr (FMT == FMT_F16X3) {
            acc[1][si][sj] = mfma16(af[0][si], bf[1][sj], acc[1][si][sj], FMT);
            acc[1][si][sj] = mfma16(af[1][si], bf[0][sj], acc[1][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[0][si], bf[0][sj], acc[0][si][sj], FMT);
          } else if constexpr (FMT == FMT_BF16X3) {
            acc[0][si][sj] = mfma16(af[0][si], bf[1][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[1][si], bf[0][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[0][si], bf[0][sj], acc[0][si][sj], FMT);
          } else {  // BF16X9: smallest terms first
            acc[0][si][sj] = mfma16(af[2][si], bf[2][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[1][si], bf[2][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[2][si], bf[1][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[1][si], bf[1][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[0][si], bf[2][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[2][si], bf[0][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[0][si], bf[1][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[1][si], bf[0][sj], acc[0][si][sj], FMT);
            acc[0][si][sj] = mfma16(af[0][si], bf[0][sj], acc[0][si][sj], FMT);
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

  f32x4 acc[NP][TL::NACC][SUB][SUB];
#pragma unroll
  for (int u = 0; u < NP; ++u)
#pragma unroll
    for (int q = 0; q < TL::NACC; ++q)
#pragma unroll
      for (int a = 0; a < SUB; ++a)
#pragma unroll
        for (int b = 0; b < SUB; ++b) acc[u][q][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

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
#define PSDP_MMA(buf)                                                                          \
  {                                                                                            \
    const E* base_ = smem + (buf) * TL::STAGE;                                                 \
    mma_slab<FMT, SUB>(base_, base_ + P * TILE, arow, brow, lane, acc[0]);                     \
    if constexpr (AVG) mma_slab<FMT, SUB>(base_ + 2 * P * TILE, base_ + 3 * P * TILE, arow, brow, lane, acc[NP - 1]); \
  }
  PSDP_GLOAD(0)
  PSDP_LSTORE(0)
  __syncthreads();
  for (int kt = 0; kt < KT - 1; ++kt) {
    PSDP_GLOAD((kt + 1) * BK)
    PSDP_MMA(kt & 1)
    PSDP_LSTORE((kt + 1) & 1)
    __syncthreads();
  }
  PSDP_MMA((KT - 1) & 1)
  __syncthreads();
#undef PSDP_GLOAD
#undef PSDP_LSTORE
#undef PSDP_MMA

  // ---------------------------------------------------------------- epilogue
  // Accumulators go to LDS with compile-time indices only (they stay in
  // registers), then the tile is written out by epilogue_from_lds.
  float* Ct = (float*)smem;
  constexpr int CS = TL::BT + 1;
#pragma unroll
  for (int si = 0; si < SUB; ++si)
#pragma unroll
    for (int sj = 0; sj < SUB; ++sj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float v = acc[0][0][si][sj][e];
        if constexpr (TL::NACC == 2) v = fmaf(acc[0][1][si][sj][e], F16X3_LO_INV, v);
        if constexpr (AVG) {
          float w = acc[NP - 1][0][si][sj][e];
          if constexpr (TL::NACC == 2) w = fmaf(acc[NP - 1][1][si][sj][e], F16X3_LO_INV, w);
          v = 0.5f * (v + w);
        }
        const int r = wr * TL::WT + si * 16 + (lane >> 4) * 4 + e;
        const int c = wc * TL::WT + sj * 16 + (lane & 15);
        Ct[r * CS + c] = v;
      }
  __syncthreads();
  epilogue_from_lds<TL::BT>(d, bi, bj, Ct);
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
    const ProdDesc d = descs[tr.mat];
    sym_tile<FMT, SUB, AVG>(d, tr.bi, tr.bj, smem);
    __syncthreads();
  }
}

template <int FMT, int SUB, bool AVG>
static hipError_t launch_one(const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s) {
  hipLaunchKernelGGL((sym_prod_kernel<FMT, SUB, AVG>), dim3(ntiles), dim3(256), 0, s, d, tiles);
  return hipGetLastError();
}

template <int FMT, int SUB, bool AVG>
static hipError_t launch_grouped(const ProdDesc* descs, const TileRef* tiles, int ntiles, int grid,
                                 hipStream_t s) {
  hipLaunchKernelGGL((sym_prod_grouped_kernel<FMT, SUB, AVG>), dim3(grid), dim3(256), 0, s, descs, tiles,
                     ntiles);
  return hipGetLastError();
}

int sym_tile_size(int sub) { return 32 * sub; }

// avg selects the averaged (Jordan) product; it exists for the 1-piece
// 16-bit formats, whose operand rounding makes C visibly non-symmetric.
hipError_t launch_sym_product(int fmt, int sub, int avg, const ProdDesc& d, const TileRef* tiles, int ntiles,
                              hipStream_t s) {
#define PSDP_CASE(F)                                                    \
  case F:                                                               \
    return sub == 4 ? launch_one<F, 4, false>(d, tiles, ntiles, s)      \
                    : launch_one<F, 2, false>(d, tiles, ntiles, s);
  if (avg) {
    if (fmt == FMT_F16)
      return sub == 4 ? launch_one<FMT_F16, 4, true>(d, tiles, ntiles, s)
                      : launch_one<FMT_F16, 2, true>(d, tiles, ntiles, s);
    if (fmt == FMT_BF16)
      return sub == 4 ? launch_one<FMT_BF16, 4, true>(d, tiles, ntiles, s)
                      : launch_one<FMT_BF16, 2, true>(d, tiles, ntiles, s);
  }
  switch (fmt) {
    PSDP_CASE(FMT_F32)
    PSDP_CASE(FMT_F16)
    PSDP_CASE(FMT_F16X3)
    PSDP_CASE(FMT_BF16)
    PSDP_CASE(FMT_BF16X3)
    PSDP_CASE(FMT_BF16X9)
    default:
      return hipErrorInvalidValue;
  }
#undef PSDP_CASE
}

hipError_t launch_sym_product_grouped(int fmt, int sub, int avg, const ProdDesc* descs, const TileRef* tiles,
                                      int ntiles, int grid, hipStream_t s) {
#define PSDP_CASE(F)                                                             \
  case F:                                                                        \
    return sub == 4 ? launch_grouped<F, 4, false>(descs, tiles, ntiles, grid, s) \
                    : launch_grouped<F, 2, false>(descs, tiles, ntiles, grid, s);
  if (avg) {
    if (fmt == FMT_F16)
      return sub == 4 ? launch_grouped<FMT_F16, 4, true>(descs, tiles, ntiles, grid, s)
                      : launch_grouped<FMT_F16, 2, true>(descs, tiles, ntiles, grid, s);
    if (fmt == FMT_BF16)
      return sub == 4 ? launch_grouped<FMT_BF16, 4, true>(descs, tiles, ntiles, grid, s)
                      : launch_grouped<FMT_BF16, 2, true>(descs, tiles, ntiles, grid, s);
  }
  switch (fmt) {
    PSDP_CASE(FMT_F32)
    PSDP_CASE(FMT_F16)
    PSDP_CASE(FMT_F16X3)
    PSDP_CASE(FMT_BF16)
    PSDP_CASE(FMT_BF16X3)
    PSDP_CASE(FMT_BF16X9)
    default:
      return hipErrorInvalidValue;
  }
#undef PSDP_CASE
}

}  // namespace psdp
