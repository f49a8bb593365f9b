// Device-resident Lanczos spectral-norm bound (approximate_two_norm).
//
// Restates src/lanczos.cu:41-293 with every scalar kept in device memory so
// that the whole bound is a fixed sequence of launches (no host round trips,
// graph capturable):
//   * Lanczos on A^T A = A A (A symmetric), no reorthogonalisation
//     (src/lanczos.cu:116-169), the break rule of :145 and :152-159 evaluated
//     on the device (except that an exact invariant subspace, or the last
//     iteration of an n x n matrix with n <= max_iter, keeps its iteration,
//     see step_body); after a break the remaining launches are no-ops;
//   * T = tridiag(alpha, beta) (src/lanczos.cu:16-38), its largest eigenpair
//     by multisection + inverse iteration in FP64 (replaces cusolverDnDsyevd
//     at :206-214; T is built from zeros, unlike the unzeroed buffer there);
//   * y = V u (:242), r = A y - theta y (the reference's :255 copies ry into
//     w1, so its residual is A y - theta y), lo = sqrt(theta) (:264),
//     up = sqrt(theta + ||r||) (:270); nb_iter == 0 gives lo = 0, up = 1
//     (:171-184).
// The matrix-vector products are bandwidth-bound.  On the float32 matrix of
// the projections they read only the tiles of the upper triangle (half the
// bytes of a full GEMV; FP64 accumulation, wave-shuffle reductions, partial
// sums combined in a fixed order by a second kernel); the float64 drop-in
// streams whole rows (one wave per RPW rows, 16-byte loads).  Vector updates
// run in one 1024-thread workgroup per matrix.  Every kernel has a
// single-matrix form and a batched form (one launch for many matrices, a flat
// work list or blockIdx.y selects the matrix) sharing the same body, so a
// matrix gets bitwise the same bound alone and in a batch.
#include "psdp_device.h"

namespace psdp {

constexpr int GEMV_RPW = 4;                    // rows per wave
constexpr int GEMV_ROWS_PER_BLOCK = 4 * GEMV_RPW;

// ------------------------------------------------------------------ GEMV
// y[r] = sum_k A[r*ld + k] x[k] for the rows of this workgroup (A symmetric:
// a row of the row-major view is a column of the column-major matrix).
template <typename TA, int VEC>
__device__ __forceinline__ void gemv_body(const TA* __restrict__ A, int64_t ld, int m, const double* __restrict__ x,
                                          double* __restrict__ y, int block) {
  const int lane = threadIdx.x & 63;
  const int r0 = (block * 4 + (threadIdx.x >> 6)) * GEMV_RPW;
  if (r0 >= m) return;
  double acc[GEMV_RPW];
#pragma unroll
  for (int r = 0; r < GEMV_RPW; ++r) acc[r] = 0.0;
  const auto Ag = gbl(A);
  const auto xg = gbl(x);
  // KU chunks of 64 VEC columns per pass: all their loads are issued before
  // the first FMA, so a wave keeps KU x RPW 16-byte loads in flight.  Each
  // accumulator still sees its terms in ascending k (the sum is the same
  // for any KU).
  constexpr int KU = 4;
  for (int k0 = lane * VEC; k0 < m; k0 += 64 * VEC * KU) {
    TA av[KU][GEMV_RPW][VEC];
    double xv[KU][VEC];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int k = k0 + u * 64 * VEC;
      if (k < m) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) xv[u][e] = xg[k + e];
#pragma unroll
        for (int r = 0; r < GEMV_RPW; ++r) {
          if (r0 + r < m) {
            const auto row = Ag + (int64_t)(r0 + r) * ld + k;
            if constexpr (VEC * sizeof(TA) == 16) {
              typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
              *(u32x4v*)av[u][r] = *(const __attribute__((address_space(1))) u32x4v*)row;
            } else {
#pragma unroll
              for (int e = 0; e < VEC; ++e) av[u][r][e] = row[e];
            }
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      if (k0 + u * 64 * VEC < m) {
#pragma unroll
        for (int r = 0; r < GEMV_RPW; ++r)
          if (r0 + r < m) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[r] = fma((double)av[u][r][e], xv[u][e], acc[r]);
          }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < GEMV_RPW; ++r) {
    const double s = wave_sum(acc[r]);
    if (lane == 0 && r0 + r < m) gbl(y)[r0 + r] = s;
  }
}

template <typename TA, int VEC>
__global__ __launch_bounds__(256) void gemv_sym_kernel(const TA* __restrict__ A, int64_t ld, int m,
                                                       const double* __restrict__ x, double* __restrict__ y,
                                                       const int* __restrict__ skip) {
  if (skip && *skip) return;
  gemv_body<TA, VEC>(A, ld, m, x, y, blockIdx.x);
}

// ------------------------------------------------------------------ symmetric GEMV
// y = A x for a symmetric float32 m x m matrix (m % 4 == 0, ld % 4 == 0,
// 16-byte aligned rows) that reads only the upper triangle's tiles.  Tile
// (I, J), I <= J, of T x T entries (T = SYMV_TILE, clipped at m) contributes
//   A_IJ x_J   to the rows of block I:  P[J][I T + r]
//   A_IJ^T x_I to the rows of block J:  P[I][J T + c]   (I < J only)
// so every P[t][i] (t < ceil(m / T), i < m) is written exactly once and
// y[i] = sum_t P[t][i] in ascending t.  All sums have a fixed order: y is
// deterministic and the same for any number of zero rows padding the matrix.
//
// In a tile, wave w takes rows w + 4 j (64 rows), lane l columns 4l .. 4l+3
// (one 16-byte load per row).  A row's direct sum is spread over the 64
// lanes: eight rows at a time are reduced by a butterfly that halves the
// values each lane carries at every exchange (4 + 2 + 1 + 3 shuffles for 8
// rows instead of 6 per row).  The transposed sums stay in the lane
// (its 4 columns over the wave's rows, x of the rows broadcast from LDS) and
// the 4 waves combine them in LDS.
__device__ __forceinline__ void symv_tile_body(const float* __restrict__ A, int64_t ld, int m,
                                               const double* __restrict__ x, double* __restrict__ P, int I,
                                               int J) {
  constexpr int T = SYMV_TILE;
  static_assert(T == 256, "4 waves x 64 rows, 64 lanes x 4 columns");
  typedef float f32x4v __attribute__((ext_vector_type(4)));
  __shared__ double red[4 * T];
  __shared__ double xs[T];  // x of the tile's rows
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = I * T, c0 = J * T;
  const int c = c0 + 4 * lane;
  const bool cin = c < m;  // m % 4 == 0: the lane's 4 columns are all inside or all outside
  const auto xg = gbl(x);
  const auto Ag = gbl(A);
  {
    const int r = r0 + (int)threadIdx.x;
    xs[threadIdx.x] = r < m ? xg[r] : 0.0;
  }
  double xc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) xc[e] = cin ? xg[c + e] : 0.0;
  __syncthreads();
  double z[4] = {0.0, 0.0, 0.0, 0.0};
  const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0, h3 = (lane & 8) != 0;
  const int kl = (h5 ? 4 : 0) + (h4 ? 2 : 0) + (h3 ? 1 : 0);  // row of the batch this lane ends with

  // batch b: rows w + 4 (8 b + k), k < 8
  // Unconditional loads (entry 0 stands in for rows/columns outside m): with
  // each load behind a branch the compiler waited for all loads (vmcnt(0)),
  // the next batch's included; now it waits per load.  (Measured: n = 16384
  // unchanged at 0.116 ms per tile launch; the 256-block batch 7.36 against
  // 7.45 ms.)
  auto load = [&](int b, f32x4v (&dst)[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int r = r0 + w + 4 * (8 * b + k);
      const bool ok = cin && r < m;
      const f32x4v v = *(const __attribute__((address_space(1))) f32x4v*)(Ag + (ok ? (int64_t)r * ld + c : 0));
      dst[k] = ok ? v : f32x4v{0.f, 0.f, 0.f, 0.f};
    }
  };
  auto process = [&](int b, const f32x4v (&cur)[8]) {
    double p[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double a0 = cur[k].x, a1 = cur[k].y, a2 = cur[k].z, a3 = cur[k].w;
      double sk = a0 * xc[0];
      sk = fma(a1, xc[1], sk);
      sk = fma(a2, xc[2], sk);
      sk = fma(a3, xc[3], sk);
      p[k] = sk;
      const double xr = xs[w + 4 * (8 * b + k)];  // same address in every lane: broadcast
      z[0] = fma(a0, xr, z[0]);
      z[1] = fma(a1, xr, z[1]);
      z[2] = fma(a2, xr, z[2]);
      z[3] = fma(a3, xr, z[3]);
    }
    // butterfly: after the exchanges across 32, 16 and 8 lanes, lane l holds
    // the partial of row kl = 4 b5 + 2 b4 + b3 (bits of l) over 8 lanes
    double q4[4], q2[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double send = h5 ? p[j] : p[j + 4];
      const double keep = h5 ? p[j + 4] : p[j];
      q4[j] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double send = h4 ? q4[j] : q4[j + 2];
      const double keep = h4 ? q4[j + 2] : q4[j];
      q2[j] = keep + __shfl_xor(send, 16);
    }
    const double send = h3 ? q2[0] : q2[1];
    const double keep = h3 ? q2[1] : q2[0];
    double t = keep + __shfl_xor(send, 8);
    t += __shfl_xor(t, 4);
    t += __shfl_xor(t, 2);
    t += __shfl_xor(t, 1);
    const int r = r0 + w + 4 * (8 * b + kl);
    if ((lane & 7) == 0 && r < m) gbl(P)[(int64_t)J * m + r] = t;
  };
  f32x4v a0[8], a1[8];
  load(0, a0);
#pragma unroll 1
  for (int b = 0; b < 8; b += 2) {  // the loads of the next batch are in flight during each batch's FMAs
    load(b + 1, a1);
    process(b, a0);
    if (b + 2 < 8) load(b + 2, a0);
    process(b + 1, a1);
  }
  if (I != J) {  // uniform over the workgroup
#pragma unroll
    for (int e = 0; e < 4; ++e) red[w * T + 4 * lane + e] = z[e];
    __syncthreads();
    const int i = threadIdx.x;
    const int col = c0 + i;
    if (col < m) gbl(P)[(int64_t)I * m + col] = ((red[i] + red[T + i]) + red[2 * T + i]) + red[3 * T + i];
  }
}

// y[i] = sum_t P[t][i] for the SYMV_TILE rows of one chunk.
__device__ __forceinline__ void symv_reduce_body(const double* __restrict__ P, int m, double* __restrict__ y,
                                                 int chunk) {
  const int i = chunk * SYMV_TILE + (int)threadIdx.x;
  if (i >= m) return;
  const int nt = (m + SYMV_TILE - 1) / SYMV_TILE;
  const auto Pg = gbl(P);
  double s = 0.0;
#pragma unroll 8
  for (int t = 0; t < nt; ++t) s += Pg[(int64_t)t * m + i];
  gbl(y)[i] = s;
}

// Tile L of the upper triangle in column order: (0,0), (0,1), (1,1), (0,2), ...
__device__ __forceinline__ void tri_decode(int L, int& I, int& J) {
  int j = (int)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= L) ++j;
  while (j * (j + 1) / 2 > L) --j;
  J = j;
  I = L - j * (j + 1) / 2;
}

__global__ __launch_bounds__(256) void symv_tile_kernel(const float* __restrict__ A, int64_t ld, int m,
                                                        const double* __restrict__ x, double* __restrict__ P,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  int I, J;
  tri_decode((int)blockIdx.x, I, J);
  symv_tile_body(A, ld, m, x, P, I, J);
}

__global__ __launch_bounds__(256) void symv_reduce_kernel(const double* __restrict__ P, int m,
                                                          double* __restrict__ y, const int* __restrict__ skip) {
  if (skip && *skip) return;
  symv_reduce_body(P, m, y, (int)blockIdx.x);
}

// which: 0 -> w1 = A q, 1 -> w = A w1, 2 -> ay = A y, 3 -> w1 = A ay (2 and 3
// skip unless there is a Ritz pair; 3 only for the tight bound).
// Flat work lists over every matrix: tiles[g] = {matrix, I, J}, chunks[g] =
// {matrix, row chunk}, so no workgroup idles on a small matrix.
__global__ __launch_bounds__(256) void symv_tile_batched(const LanczosJob* __restrict__ jobs,
                                                         const TileRef* __restrict__ tiles, int which) {
  const TileRef t = tiles[blockIdx.x];
  const LanczosJob& J = jobs[t.mat];
  if (which < 2 ? J.st->done : J.st->skip_ritz) return;
  const double* x = which == 0 ? J.v.q : (which == 1 ? J.v.w1 : (which == 2 ? J.v.y : J.v.ay));
  symv_tile_body(J.A, J.ld, J.m, x, J.v.part, t.bi, t.bj);
}

__global__ __launch_bounds__(256) void symv_reduce_batched(const LanczosJob* __restrict__ jobs,
                                                           const int2* __restrict__ chunks, int which) {
  const int2 ch = chunks[blockIdx.x];
  const LanczosJob& J = jobs[ch.x];
  if (which < 2 ? J.st->done : J.st->skip_ritz) return;
  double* y = which == 0 ? J.v.w1 : (which == 1 ? J.v.w : (which == 2 ? J.v.ay : J.v.w1));
  symv_reduce_body(J.v.part, J.m, y, ch.y);
}

// ------------------------------------------------------------------ start

// q_i ~ U(0,1] for i < n (zero in the padding), q <- q/||q||, V(:,0) = q,
// vold = 0 (src/lanczos.cu:69,101-112).  The start vector is a seeded
// counter-based stream instead of cuRAND XORWOW with a clock seed.
__device__ void init_body(LanczosState* st, const LanczosVecs& v, int n, int m, uint64_t seed) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double qi = 0.0;
    if (i < n)
      qi = uniform01(seed, (uint64_t)i);
    v.q[i] = qi;
    s += qi * qi;
  }
  const double nrm = sqrt(block_sum(s, red));
  const double inv = nrm != 0.0 ? 1.0 / nrm : 1.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double qi = v.q[i] * inv;
    v.q[i] = qi;
    v.V[i] = qi;
    v.vold[i] = 0.0;
  }
  if (threadIdx.x == 0) {
    st->beta_old = 0.0;
    st->nb = 0;
    st->done = 0;
    st->skip_ritz = 0;
    st->nonfinite = 0;
    st->out_nonfinite = 0;
    st->theta = 0.0;
    st->rnorm = 0.0;
    st->lo = 0.0;
    st->up = 1.0;
    st->scale = 1.0;
    st->inv_scale = 1.0;
  }
}

// One Lanczos iteration after w = A (A q) (src/lanczos.cu:130-168); n is the
// logical size of the matrix.  With parts = true, w is first formed here from
// the symmetric GEMV's partial sums (w[i] = sum_t P[t][i] in ascending t, the
// sum symv_reduce_body forms, so w is bitwise the same) instead of by a
// reduction launch: the batched bound saves one launch per iteration.
template <bool parts>
__device__ void step_body(LanczosState* st, const LanczosVecs& v, int k, int n, int m, int max_iter, double tol) {
  __shared__ double red[16];
  if (st->done || k >= max_iter) return;
  double s = 0.0;
  if constexpr (parts) {
    const int nt = (m + SYMV_TILE - 1) / SYMV_TILE;
    const auto Pg = gbl(v.part);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double wi = 0.0;
      for (int t = 0; t < nt; ++t) wi += Pg[(int64_t)t * m + i];
      v.w[i] = wi;
      s += v.q[i] * wi;
    }
  } else {
    for (int i = threadIdx.x; i < m; i += blockDim.x) s += v.q[i] * v.w[i];
  }
  const double alpha = block_sum(s, red);
  const double mbeta_old = -st->beta_old;
  s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double wi = fma(-alpha, v.q[i], v.w[i]);
    wi = fma(mbeta_old, v.vold[i], wi);
    v.w[i] = wi;
    s += wi * wi;
  }
  const double beta = sqrt(block_sum(s, red));
  const bool bad = !isfinite(alpha) || !isfinite(beta);
  // src/lanczos.cu:145 drops iteration k on convergence.  On an exact
  // invariant subspace (beta == 0, src/lanczos.cu:152-159) the reference also
  // drops it, so a start vector that is an eigenvector (any 1x1 matrix) is
  // taken for the zero matrix and gets up = 1; here that iteration is kept
  // (its Ritz value is exact) unless alpha == 0 (the zero matrix itself).
  // Likewise iteration n - 1 of an n x n matrix (n <= max_iter): the Krylov
  // space is then the whole space, beta is rounding noise, and dropping the
  // iteration leaves theta from an (n-1)-dimensional space, which can fall
  // below ||A||^2 by more than the residual term adds back (the bound then
  // undercuts ||A|| and the filter chain diverges on a 3 x 3 input); kept,
  // theta is the exact largest eigenvalue of A A.
  const bool conv = beta <= tol * alpha && k > 1;
  const bool full = conv && k == n - 1 && alpha > 0.0;
  const bool drop = bad || (conv && !full) || (beta == 0.0 && !(alpha > 0.0));
  const bool exact = !drop && (beta == 0.0 || full);
  const bool stop = drop || exact;
  if (!stop) {
    const double inv = 1.0 / beta;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const double qi = v.w[i] * inv;
      v.vold[i] = v.q[i];
      v.q[i] = qi;
      if (k < max_iter - 1) v.V[(int64_t)(k + 1) * m + i] = qi;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st->alpha[k] = alpha;
    st->beta[k] = beta;
    if (bad) st->nonfinite = 1;
    if (drop) {
      st->done = 1;
    } else if (exact) {
      st->nb = k + 1;
      st->done = 1;
    } else {
      st->beta_old = beta;
      st->nb = k + 1;
      if (k == max_iter - 1) st->done = 1;
    }
  }
}

// Largest eigenpair of T = tridiag(alpha[0..nb), beta[0..nb-1)) in FP64,
// one wave (replaces cusolverDnDsyevd on T, src/lanczos.cu:206-214):
//   * the eigenvalue by multisection on Sturm counts: the 64 lanes evaluate
//     the number of eigenvalues below 64 points of the bracket at once, so
//     each round shrinks it 65x (about 9 rounds to full precision);
//   * the eigenvector by inverse iteration on T - theta I with a partially
//     pivoted tridiagonal LU (the LAPACK dgttrf/dgtts2 scheme), 3 solves.
__device__ __forceinline__ int sturm_count(const double* a, const double* b2, int nb, double x, double pivmin) {
  int cnt = 0;
  double d = 1.0;
  for (int i = 0; i < nb; ++i) {
    d = (a[i] - x) - (i ? b2[i - 1] / d : 0.0);
    if (fabs(d) < pivmin) d = -pivmin;
    cnt += d < 0.0;
  }
  return cnt;  // eigenvalues below x
}

__device__ void ritz_body(LanczosState* st) {
  __shared__ double a[LANCZOS_MAX_ITER], b[LANCZOS_MAX_ITER], b2[LANCZOS_MAX_ITER];
  // work arrays of the (serial) inverse iteration, in LDS
  __shared__ double u[LANCZOS_MAX_ITER], dl[LANCZOS_MAX_ITER], dd[LANCZOS_MAX_ITER], du[LANCZOS_MAX_ITER],
      du2[LANCZOS_MAX_ITER];
  __shared__ int piv[LANCZOS_MAX_ITER];
  const int lane = threadIdx.x;
  const int nb = st->nb;
  if (nb == 0 || st->nonfinite) {
    if (lane == 0) {
      st->skip_ritz = 1;
      if (st->nonfinite) {
        st->lo = st->up = __builtin_nan("");
        st->scale = st->inv_scale = __builtin_nan("");
      } else {
        st->lo = 0.0;
        st->up = 1.0;
        st->scale = 1.0;
        st->inv_scale = 1.0;
      }
    }
    return;
  }
  if (lane < nb) {
    a[lane] = st->alpha[lane];
    b[lane] = lane + 1 < nb ? st->beta[lane] : 0.0;
    b2[lane] = b[lane] * b[lane];
  }
  __syncthreads();
  // Gershgorin bracket and the pivot floor of the Sturm recurrence
  double glo = 1e308, ghi = -1e308, bmax = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double r = fabs(b[i]) + (i ? fabs(b[i - 1]) : 0.0);
    glo = fmin(glo, a[i] - r);
    ghi = fmax(ghi, a[i] + r);
    bmax = fmax(bmax, b2[i]);
  }
  const double pivmin = 1e-290 * fmax(1.0, bmax);
  const double wid = fmax(ghi - glo, 1e-300);
  double L = glo - 1e-12 * wid - 1e-300, H = ghi + 1e-12 * wid + 1e-300;  // count(L) = 0, count(H) = nb
  for (int round = 0; round < 64; ++round) {
    if (H - L <= 4.0 * 2.2204460492503131e-16 * fmax(fabs(L), fabs(H)) || H - L <= 1e-300) break;
    const double x = L + (H - L) * (double)(lane + 1) / 65.0;
    const bool top = sturm_count(a, b2, nb, x, pivmin) == nb;  // every eigenvalue below x
    const unsigned long long m = __ballot(top);
    if (m) {
      const int j = __ffsll((long long)m) - 1;                // first point above the largest
      const double xj = L + (H - L) * (double)(j + 1) / 65.0;
      const double xl = j ? L + (H - L) * (double)j / 65.0 : L;
      H = xj;
      L = xl;
    } else {
      L = L + (H - L) * 64.0 / 65.0;
    }
  }
  const double theta = 0.5 * (L + H);
  if (lane == 0) {
    const double tiny = 2.2204460492503131e-16 * fmax(fabs(glo), fabs(ghi)) + 1e-300;
    // LU of T - theta I with partial pivoting (rows i, i+1)
    for (int i = 0; i < nb; ++i) {
      dd[i] = a[i] - theta;
      dl[i] = b[i];
      du[i] = b[i];
      du2[i] = 0.0;
    }
    for (int i = 0; i + 1 < nb; ++i) {
      if (fabs(dd[i]) >= fabs(dl[i])) {
        if (dd[i] == 0.0) dd[i] = tiny;
        const double f = dl[i] / dd[i];
        dl[i] = f;
        dd[i + 1] -= f * du[i];
        piv[i] = i;
      } else {
        const double f = dd[i] / dl[i];
        dd[i] = dl[i];
        dl[i] = f;
        const double t = du[i];
        du[i] = dd[i + 1];
        dd[i + 1] = t - f * dd[i + 1];
        if (i + 2 < nb) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du[i + 1];
        }
        piv[i] = i + 1;
      }
    }
    if (fabs(dd[nb - 1]) < tiny) dd[nb - 1] = tiny;
    for (int i = 0; i < nb; ++i) u[i] = 1.0;
    for (int it = 0; it < 3; ++it) {
      for (int i = 0; i + 1 < nb; ++i) {  // L solve
        if (piv[i] == i) {
          u[i + 1] -= dl[i] * u[i];
        } else {
          const double t = u[i];
          u[i] = u[i + 1];
          u[i + 1] = t - dl[i] * u[i];
        }
      }
      for (int i = nb - 1; i >= 0; --i) {  // U solve
        double r = u[i];
        if (i + 1 < nb) r -= du[i] * u[i + 1];
        if (i + 2 < nb) r -= du2[i] * u[i + 2];
        u[i] = r / dd[i];
      }
      double nrm = 0.0;
      for (int i = 0; i < nb; ++i) nrm = fmax(nrm, fabs(u[i]));
      for (int i = 0; i < nb; ++i) u[i] /= nrm;
    }
    double s2 = 0.0;
    for (int i = 0; i < nb; ++i) s2 += u[i] * u[i];
    const double inv = 1.0 / sqrt(s2);
    for (int i = 0; i < nb; ++i) st->u[i] = u[i] * inv;
    st->theta = theta;
  }
}

// y = V(:, 0:nb) u  (src/lanczos.cu:242)
__device__ __forceinline__ void ritz_vector_body(const LanczosState* st, const LanczosVecs& v, int m, int i) {
  if (st->skip_ritz || i >= m) return;
  const int nb = st->nb;
  double s = 0.0;
  for (int j = 0; j < nb; ++j) s = fma(v.V[(int64_t)j * m + i], st->u[j], s);
  v.y[i] = s;
}

// r = A y - theta y; lo = sqrt(theta); up = sqrt(theta + ||r||)
// (src/lanczos.cu:258-270); s = up > 0 ? up : 1 (src/composite_FP32.cu:182).
// tight: r = A (A y) - theta y (v.w1 holds A A y), the residual of the Ritz
// pair of A^T A itself, so that up = sqrt(theta + ||r||) exceeds ||A||_2 by
// the Ritz pair's own error instead of by |lambda - lambda^2|.
// With parts = true, the last GEMV's result (ay, or w1 for the tight bound)
// is first formed here from its partial sums, as in step_body.
template <bool parts>
__device__ void bound_body(LanczosState* st, const LanczosVecs& v, int m, double margin, int tight) {
  __shared__ double red[16];
  if (st->skip_ritz) return;
  const double th = st->theta;
  double* ry = tight ? v.w1 : v.ay;
  double s = 0.0;
  const int nt = (m + SYMV_TILE - 1) / SYMV_TILE;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double ri;
    if constexpr (parts) {
      const auto Pg = gbl(v.part);
      ri = 0.0;
      for (int t = 0; t < nt; ++t) ri += Pg[(int64_t)t * m + i];
      ry[i] = ri;
    } else {
      ri = ry[i];
    }
    const double r = fma(-th, v.y[i], ri);
    s += r * r;
  }
  const double rn = sqrt(block_sum(s, red));
  if (threadIdx.x == 0) {
    // theta >= 0 in exact arithmetic (Ritz value of A^T A); clamp rounding
    const double thc = th > 0.0 ? th : 0.0;
    const double up = sqrt(thc + rn);
    st->rnorm = rn;
    st->lo = sqrt(thc);
    st->up = up;
    const double sc = up > 0.0 ? up * margin : 1.0;
    st->scale = sc;
    st->inv_scale = 1.0 / sc;
    if (!isfinite(up)) st->nonfinite = 1;
  }
}

// ------------------------------------------------------------------ single
__global__ __launch_bounds__(1024) void lanczos_init_kernel(LanczosState* st, LanczosVecs v, int n, int m,
                                                            uint64_t seed) {
  init_body(st, v, n, m, seed);
}
__global__ __launch_bounds__(1024) void lanczos_step_kernel(LanczosState* st, LanczosVecs v, int k, int n, int m,
                                                            int max_iter, double tol) {
  step_body<false>(st, v, k, n, m, max_iter, tol);
}
__global__ __launch_bounds__(64) void lanczos_ritz_kernel(LanczosState* st) { ritz_body(st); }
__global__ __launch_bounds__(256) void ritz_vector_kernel(const LanczosState* st, LanczosVecs v, int m) {
  ritz_vector_body(st, v, m, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ __launch_bounds__(1024) void lanczos_bound_kernel(LanczosState* st, LanczosVecs v, int m,
                                                             double margin, int tight) {
  bound_body<false>(st, v, m, margin, tight);
}

// ------------------------------------------------------------------ batched
__global__ __launch_bounds__(1024) void lanczos_init_batched(const LanczosJob* jobs, uint64_t seed) {
  const LanczosJob& J = jobs[blockIdx.x];
  init_body(J.st, J.v, J.n, J.m, seed);
}
__global__ __launch_bounds__(1024) void lanczos_step_batched(const LanczosJob* jobs, int k, double tol) {
  const LanczosJob& J = jobs[blockIdx.x];
  step_body<true>(J.st, J.v, k, J.n, J.m, J.max_iter, tol);
}
__global__ __launch_bounds__(64) void lanczos_ritz_batched(const LanczosJob* jobs) { ritz_body(jobs[blockIdx.x].st); }
__global__ __launch_bounds__(256) void ritz_vector_batched(const LanczosJob* jobs) {
  const LanczosJob& J = jobs[blockIdx.y];
  ritz_vector_body(J.st, J.v, J.m, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ __launch_bounds__(1024) void lanczos_bound_batched(const LanczosJob* jobs, double margin, int tight) {
  const LanczosJob& J = jobs[blockIdx.x];
  bound_body<true>(J.st, J.v, J.m, margin, tight);
}

// Fixed scale (no bound): used by the non-auto-scaled entry points.
__global__ void set_scale_kernel(LanczosState* st, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->scale = scale;
    st->inv_scale = 1.0 / scale;
    st->nonfinite = 0;
    st->out_nonfinite = 0;
  }
}

template <typename TA>
static void launch_gemv(const TA* A, int64_t ld, int m, const double* x, double* y, const int* skip,
                        hipStream_t s) {
  const int grid = (m + GEMV_ROWS_PER_BLOCK - 1) / GEMV_ROWS_PER_BLOCK;
  constexpr int VEC = 16 / (int)sizeof(TA);
  const bool aligned = (m % VEC == 0) && (ld % VEC == 0) && (((uintptr_t)A) % 16 == 0);
  if (aligned)
    hipLaunchKernelGGL((gemv_sym_kernel<TA, VEC>), dim3(grid), dim3(256), 0, s, A, ld, m, x, y, skip);
  else
    hipLaunchKernelGGL((gemv_sym_kernel<TA, 1>), dim3(grid), dim3(256), 0, s, A, ld, m, x, y, skip);
}

// y = A x on the upper triangle of a float32 matrix (P: ceil(m/T) x m doubles).
static void launch_symv(const float* A, int64_t ld, int m, const double* x, double* y, double* P, const int* skip,
                        hipStream_t s) {
  const int nt = (m + SYMV_TILE - 1) / SYMV_TILE;
  hipLaunchKernelGGL(symv_tile_kernel, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, A, ld, m, x, P, skip);
  hipLaunchKernelGGL(symv_reduce_kernel, dim3(nt), dim3(256), 0, s, P, m, y, skip);
}

// Enqueues the whole bound.  A: symmetric m x m (row-major view = column-major),
// logical size n <= m (entries beyond n are zero).  Vectors have length m.
// A float32 A needs m % 4 == 0, ld % 4 == 0, 16-byte alignment and v.part.
hipError_t launch_lanczos(const void* A, int a_f64, int64_t ld, int n, int m, int max_iter, double tol,
                          uint64_t seed, double margin, int tight, LanczosState* st, const LanczosVecs& v,
                          hipStream_t s) {
  if (!a_f64 && (m % 4 != 0 || ld % 4 != 0 || ((uintptr_t)A) % 16 != 0 || !v.part)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lanczos_init_kernel, dim3(1), dim3(1024), 0, s, st, v, n, m, seed);
  for (int k = 0; k < max_iter; ++k) {
    if (a_f64) {
      launch_gemv((const double*)A, ld, m, v.q, v.w1, &st->done, s);
      launch_gemv((const double*)A, ld, m, v.w1, v.w, &st->done, s);
    } else {
      launch_symv((const float*)A, ld, m, v.q, v.w1, v.part, &st->done, s);
      launch_symv((const float*)A, ld, m, v.w1, v.w, v.part, &st->done, s);
    }
    hipLaunchKernelGGL(lanczos_step_kernel, dim3(1), dim3(1024), 0, s, st, v, k, n, m, max_iter, tol);
  }
  hipLaunchKernelGGL(lanczos_ritz_kernel, dim3(1), dim3(64), 0, s, st);
  hipLaunchKernelGGL(ritz_vector_kernel, dim3((m + 255) / 256), dim3(256), 0, s, st, v, m);
  if (a_f64)
    launch_gemv((const double*)A, ld, m, v.y, v.ay, &st->skip_ritz, s);
  else
    launch_symv((const float*)A, ld, m, v.y, v.ay, v.part, &st->skip_ritz, s);
  if (tight) {
    if (a_f64)
      launch_gemv((const double*)A, ld, m, v.ay, v.w1, &st->skip_ritz, s);
    else
      launch_symv((const float*)A, ld, m, v.ay, v.w1, v.part, &st->skip_ritz, s);
  }
  hipLaunchKernelGGL(lanczos_bound_kernel, dim3(1), dim3(1024), 0, s, st, v, m, margin, tight);
  return hipGetLastError();
}

// Batched bound over `count` float32 padded matrices (jobs in device memory;
// every job's m and ld are multiples of 4, m is at most max_m, each job has
// its partial-sum buffer v.part).  tiles: {matrix, I, J} for every upper
// triangle tile (I <= J < ceil(m / SYMV_TILE)) of every matrix; chunks:
// {matrix, c} for every SYMV_TILE-row chunk c of every matrix.
hipError_t launch_lanczos_batched(const LanczosJob* jobs, int count, int max_m, int max_iter, double tol,
                                  uint64_t seed, double margin, int tight, const TileRef* tiles, int ntiles,
                                  const int2* chunks, int nchunks, hipStream_t s) {
  const dim3 tgrid(ntiles), rgrid(nchunks);
  hipLaunchKernelGGL(lanczos_init_batched, dim3(count), dim3(1024), 0, s, jobs, seed);
  for (int k = 0; k < max_iter; ++k) {
    hipLaunchKernelGGL(symv_tile_batched, tgrid, dim3(256), 0, s, jobs, tiles, 0);
    hipLaunchKernelGGL(symv_reduce_batched, rgrid, dim3(256), 0, s, jobs, chunks, 0);
    hipLaunchKernelGGL(symv_tile_batched, tgrid, dim3(256), 0, s, jobs, tiles, 1);
    // w = sum of the partial sums: formed by the step itself
    hipLaunchKernelGGL(lanczos_step_batched, dim3(count), dim3(1024), 0, s, jobs, k, tol);
  }
  hipLaunchKernelGGL(lanczos_ritz_batched, dim3(count), dim3(64), 0, s, jobs);
  hipLaunchKernelGGL(ritz_vector_batched, dim3((max_m + 255) / 256, count), dim3(256), 0, s, jobs);
  const int last = tight ? 3 : 2;
  for (int which = 2; which <= last; ++which) {
    hipLaunchKernelGGL(symv_tile_batched, tgrid, dim3(256), 0, s, jobs, tiles, which);
    // the last result's partial sums are added by the bound kernel itself
    if (which < last) hipLaunchKernelGGL(symv_reduce_batched, rgrid, dim3(256), 0, s, jobs, chunks, which);
  }
  hipLaunchKernelGGL(lanczos_bound_batched, dim3(count), dim3(1024), 0, s, jobs, margin, tight);
  return hipGetLastError();
}

// y = A x for a symmetric float64 matrix (also used by the eigenpair Lanczos).
hipError_t launch_gemv_f64(const double* A, int64_t ld, int m, const double* x, double* y, const int* skip,
                           hipStream_t s) {
  launch_gemv(A, ld, m, x, y, skip, s);
  return hipGetLastError();
}

hipError_t launch_set_scale(LanczosState* st, double scale, hipStream_t s) {
  hipLaunchKernelGGL(set_scale_kernel, dim3(1), dim3(64), 0, s, st, scale);
  return hipGetLastError();
}

}  // namespace psdp
