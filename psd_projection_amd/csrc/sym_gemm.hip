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

template <int FMT, int SUB>
struct Tiling {
  using F = FmtTraits<FMT>;
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
  static constexpr int STAGE = 2 * P * TILE;   // X pieces then Y pieces
  static constexpr int LDS_BYTES = 2 * STAGE * ESZ;
  static_assert(CHUNKS % 256 == 0, "tile too small for 256 threads");
};

__device__ __forceinline__ f32x4 mfma16(uint4 a, uint4 b, f32x4 c, int fmt) {
  if (fmt == FMT_BF16 || fmt == FMT_BF16X3 || fmt == FMT_BF16X9)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// Writes the epilogue value w of element (i, j) and its mirror.
struct EpiWriter {
  const ProdDesc& d;
  __device__ __forceinline__ void put(int64_t idx, float w) const {
    if (d.mode == EPI_OUTPUT) return;
    if (d.c32) d.c32[idx] = w;
    if (d.out_fmt != FMT_F32) {
      uint16_t p[3];
      split_value(d.out_fmt, w, p);
      const int np = fmt_pieces(d.out_fmt);
      d.outp[0][idx] = p[0];
      if (np > 1) d.outp[1][idx] = p[1];
      if (np > 2) d.outp[2][idx] = p[2];
    }
  }
};

__device__ __forceinline__ float epi_value(const ProdDesc& d, int i, int j, float v) {
  if (d.mode == EPI_POLY || d.mode == EPI_POLY_HALF_ID) {
    const int64_t ij = (int64_t)i * d.ld + j;
    float w = d.a * d.c32[ij];
    w = fmaf(d.b, d.p3[ij], w);
    w = fmaf(d.c, v, w);
    if (d.mode == EPI_POLY_HALF_ID) w = (w + (i == j ? 1.0f : 0.0f)) * 0.5f;
    return w;
  }
  return v;
}

__device__ __forceinline__ void epi_output(const ProdDesc& d, int i, int j, float v, double s) {
  if (i >= d.n || j >= d.n) return;
  if (d.out_f64) {
    double* o = (double*)d.out;
    const double w = (double)v * s;
    o[(int64_t)i + (int64_t)j * d.out_ld] = w;
    o[(int64_t)j + (int64_t)i * d.out_ld] = w;
  } else {
    float* o = (float*)d.out;
    const float w = (float)((double)v * s);
    o[(int64_t)i + (int64_t)j * d.out_ld] = w;
    o[(int64_t)j + (int64_t)i * d.out_ld] = w;
  }
}

template <int FMT, int SUB>
__device__ __forceinline__ void sym_tile(const ProdDesc& d, int bi, int bj,
                                         typename Tiling<FMT, SUB>::E* smem) {
  using TL = Tiling<FMT, SUB>;
  using E = typename TL::E;
  constexpr int P = TL::P, LOADS = TL::LOADS, BK = TL::BK, ROW = TL::ROW, TILE = TL::TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t ld = d.ld;
  const int row0 = bi * TL::BT, col0 = bj * TL::BT;

  const E* xs[P];
  const E* ys[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    xs[p] = (const E*)d.x[p] + (int64_t)row0 * ld;
    ys[p] = (const E*)d.y[p] + (int64_t)col0 * ld;
  }
  int goff[LOADS], soff[LOADS];
#pragma unroll
  for (int i = 0; i < LOADS; ++i) {
    const int c = tid + i * 256;
    const int r = c / TL::CPR, cc = (c % TL::CPR) * TL::VEC;
    goff[i] = r * (int)ld + cc;  // r < BT, so r*ld stays far below 2^31
    soff[i] = r * ROW + cc;
  }

  uint4 rx[P][LOADS], ry[P][LOADS];
  f32x4 acc[TL::NACC][SUB][SUB];
#pragma unroll
  for (int q = 0; q < TL::NACC; ++q)
#pragma unroll
    for (int a = 0; a < SUB; ++a)
#pragma unroll
      for (int b = 0; b < SUB; ++b) acc[q][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  auto gload = [&](int k0) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int i = 0; i < LOADS; ++i) {
        rx[p][i] = *(const uint4*)(xs[p] + goff[i] + k0);
        ry[p][i] = *(const uint4*)(ys[p] + goff[i] + k0);
      }
  };
  auto lstore = [&](int buf) {
    E* base = smem + buf * TL::STAGE;
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int i = 0; i < LOADS; ++i) {
        *(uint4*)(base + p * TILE + soff[i]) = rx[p][i];
        *(uint4*)(base + (P + p) * TILE + soff[i]) = ry[p][i];
      }
  };

  const int arow = (wr * TL::WT + (lane & 15)) * ROW;
  const int brow = (wc * TL::WT + (lane & 15)) * ROW;

  auto compute = [&](int buf) {
    const E* X = smem + buf * TL::STAGE;
    const E* Y = X + P * TILE;
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
              acc[0][si][sj] =
                  __builtin_amdgcn_mfma_f32_16x16x4f32(af[si][q], bf[sj][q], acc[0][si][sj], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 32) {
        uint4 af[P][SUB], bf[P][SUB];
        const int ko = kk + (lane >> 4) * 8;
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
          for (int s = 0; s < SUB; ++s) {
            af[p][s] = *(const uint4*)(X + p * TILE + arow + s * 16 * ROW + ko);
            bf[p][s] = *(const uint4*)(Y + p * TILE + brow + s * 16 * ROW + ko);
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
  };

  const int KT = d.npad / BK;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const bool more = kt + 1 < KT;
    if (more) gload((kt + 1) * BK);
    compute(kt & 1);
    if (more) lstore((kt + 1) & 1);
    __syncthreads();
  }

  // ---------------------------------------------------------------- epilogue
  const bool diag = bi == bj;
  const double oscale = (d.mode == EPI_OUTPUT && d.scale) ? *d.scale : 1.0;
  EpiWriter wtr{d};
#pragma unroll
  for (int si = 0; si < SUB; ++si)
#pragma unroll
    for (int sj = 0; sj < SUB; ++sj) {
      const int ib = row0 + wr * TL::WT + si * 16 + (lane >> 4) * 4;
      const int j = col0 + wc * TL::WT + sj * 16 + (lane & 15);
      float w[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        w[e] = acc[0][si][sj][e];
        if constexpr (TL::NACC == 2) w[e] = fmaf(acc[1][si][sj][e], F16X3_LO_INV, w[e]);
      }
      if (d.mode == EPI_OUTPUT) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (!diag || ib + e <= j) epi_output(d, ib + e, j, w[e], oscale);
        continue;
      }
      if (diag) {
        // Only elements on or above the diagonal are read and written, so no
        // lane reads an A value that another lane's mirror store replaces.
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = ib + e;
          if (i <= j) {
            const float x = epi_value(d, i, j, w[e]);
            wtr.put((int64_t)i * ld + j, x);
            if (i != j) wtr.put((int64_t)j * ld + i, x);
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = epi_value(d, ib + e, j, w[e]);
        // primary: 4 rows, one column; mirror: one row, 4 consecutive columns
#pragma unroll
        for (int e = 0; e < 4; ++e) wtr.put((int64_t)(ib + e) * ld + j, w[e]);
        const int64_t m = (int64_t)j * ld + ib;
        if (d.c32) *(f32x4*)(d.c32 + m) = f32x4{w[0], w[1], w[2], w[3]};
        if (d.out_fmt != FMT_F32) {
          uint16_t p[4][3];
#pragma unroll
          for (int e = 0; e < 4; ++e) split_value(d.out_fmt, w[e], p[e]);
          const int np = fmt_pieces(d.out_fmt);
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            if (q < np) {
              uint2 pk;
              pk.x = (uint32_t)p[0][q] | ((uint32_t)p[1][q] << 16);
              pk.y = (uint32_t)p[2][q] | ((uint32_t)p[3][q] << 16);
              *(uint2*)(d.outp[q] + m) = pk;
            }
          }
        }
      }
    }
}

// One workgroup per upper-triangle tile of a single matrix.
template <int FMT, int SUB>
__global__ __launch_bounds__(256) void sym_prod_kernel(ProdDesc d, const TileRef* __restrict__ tiles) {
  using TL = Tiling<FMT, SUB>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[2 * TL::STAGE];
  const TileRef t = tiles[blockIdx.x];
  if (t.bi < 0) return;
  sym_tile<FMT, SUB>(d, t.bi, t.bj, smem);
}

// Persistent grouped launch: workgroups stride over a tile list that spans
// many matrices (tiles sorted by cost on the host).
template <int FMT, int SUB>
__global__ __launch_bounds__(256) void sym_prod_grouped_kernel(const ProdDesc* __restrict__ descs,
                                                               const TileRef* __restrict__ tiles,
                                                               int ntiles) {
  using TL = Tiling<FMT, SUB>;
  __shared__ __attribute__((aligned(16))) typename TL::E smem[2 * TL::STAGE];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const ProdDesc d = descs[tr.mat];
    sym_tile<FMT, SUB>(d, tr.bi, tr.bj, smem);
    __syncthreads();
  }
}

template <int FMT, int SUB>
static hipError_t launch_one(const ProdDesc& d, const TileRef* tiles, int ntiles, hipStream_t s) {
  hipLaunchKernelGGL((sym_prod_kernel<FMT, SUB>), dim3(ntiles), dim3(256), 0, s, d, tiles);
  return hipGetLastError();
}

template <int FMT, int SUB>
static hipError_t launch_grouped(const ProdDesc* descs, const TileRef* tiles, int ntiles, int grid,
                                 hipStream_t s) {
  hipLaunchKernelGGL((sym_prod_grouped_kernel<FMT, SUB>), dim3(grid), dim3(256), 0, s, descs, tiles,
                     ntiles);
  return hipGetLastError();
}

int sym_tile_size(int sub) { return 32 * sub; }

hipError_t launch_sym_product(int fmt, int sub, const ProdDesc& d, const TileRef* tiles, int ntiles,
                              hipStream_t s) {
#define PSDP_CASE(F)                                                    \
  case F:                                                               \
    return sub == 4 ? launch_one<F, 4>(d, tiles, ntiles, s)             \
                    : launch_one<F, 2>(d, tiles, ntiles, s);
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

hipError_t launch_sym_product_grouped(int fmt, int sub, const ProdDesc* descs, const TileRef* tiles,
                                      int ntiles, int grid, hipStream_t s) {
#define PSDP_CASE(F)                                                             \
  case F:                                                                        \
    return sub == 4 ? launch_grouped<F, 4>(descs, tiles, ntiles, grid, s)        \
                    : launch_grouped<F, 2>(descs, tiles, ntiles, grid, s);
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
