// Device-resident Lanczos spectral-norm bound (approximate_two_norm).
//
// Restates src/lanczos.cu:41-293 with every scalar kept in device memory so
// that the whole bound is a fixed sequence of launches (no host round trips,
// graph capturable):
//   * Lanczos on A^T A = A A (A symmetric), no reorthogonalisation
//     (src/lanczos.cu:116-169), the break rule of :145 and :152-159 evaluated
//     on the device (except that an exact invariant subspace keeps its
//     iteration, see step_body); after a break the remaining launches are
//     no-ops;
//   * T = tridiag(alpha, beta) (src/lanczos.cu:16-38), its largest eigenpair
//     by multisection + inverse iteration in FP64 (replaces cusolverDnDsyevd
//     at :206-214; T is built from zeros, unlike the unzeroed buffer there);
//   * y = V u (:242), r = A y - theta y (the reference's :255 copies ry into
//     w1, so its residual is A y - theta y), lo = sqrt(theta) (:264),
//     up = sqrt(theta + ||r||) (:270); nb_iter == 0 gives lo = 0, up = 1
//     (:171-184).
// The matrix-vector products are a bandwidth-bound GEMV (one wave streams
// RPW rows with 16-byte loads, FP64 accumulation, wave-shuffle reduction);
// vector updates run in one 1024-thread workgroup per matrix.  Every kernel
// has a single-matrix form and a batched form (one launch for many matrices,
// blockIdx.y or blockIdx.x selects the matrix) sharing the same body, so a
// matrix gets bitwise the same bound alone and in a batch.
#include "psdp_device.h"

namespace psdp {

constexpr int GEMV_RPW = 4;                    // rows per wave
constexpr int GEMV_ROWS_PER_BLOCK = 4 * GEMV_RPW;
static_assert(GEMV_ROWS_PER_BLOCK == GEMV_BATCH_ROWS, "batched GEMV row groups");

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

// which: 0 -> w1 = A q, 1 -> w = A w1, 2 -> ay = A y (skip unless Ritz pair).
// Workgroup g takes the GEMV_ROWS_PER_BLOCK rows rows[g].y of matrix rows[g].x
// (a flat list over every matrix's row groups, so no workgroup idles).
__global__ __launch_bounds__(256) void gemv_batched_kernel(const LanczosJob* __restrict__ jobs,
                                                           const int2* __restrict__ rows, int which) {
  const int2 rg = rows[blockIdx.x];
  const LanczosJob& J = jobs[rg.x];
  if (which < 2 ? J.st->done : J.st->skip_ritz) return;
  const double* x = which == 0 ? J.v.q : (which == 1 ? J.v.w1 : J.v.y);
  double* y = which == 0 ? J.v.w1 : (which == 1 ? J.v.w : J.v.ay);
  gemv_body<float, 4>(J.A, J.ld, J.m, x, y, rg.y);
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
    st->theta = 0.0;
    st->rnorm = 0.0;
    st->lo = 0.0;
    st->up = 1.0;
    st->scale = 1.0;
    st->inv_scale = 1.0;
  }
}

// One Lanczos iteration after w = A (A q) (src/lanczos.cu:130-168).
__device__ void step_body(LanczosState* st, const LanczosVecs& v, int k, int m, int max_iter, double tol) {
  __shared__ double red[16];
  if (st->done || k >= max_iter) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s += v.q[i] * v.w[i];
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
  const bool drop = bad || (beta <= tol * alpha && k > 1) || (beta == 0.0 && !(alpha > 0.0));
  const bool exact = !drop && beta == 0.0;
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
__device__ void bound_body(LanczosState* st, const LanczosVecs& v, int m, double margin) {
  __shared__ double red[16];
  if (st->skip_ritz) return;
  const double th = st->theta;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double r = fma(-th, v.y[i], v.ay[i]);
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
__global__ __launch_bounds__(1024) void lanczos_step_kernel(LanczosState* st, LanczosVecs v, int k, int m,
                                                            int max_iter, double tol) {
  step_body(st, v, k, m, max_iter, tol);
}
__global__ __launch_bounds__(64) void lanczos_ritz_kernel(LanczosState* st) { ritz_body(st); }
__global__ __launch_bounds__(256) void ritz_vector_kernel(const LanczosState* st, LanczosVecs v, int m) {
  ritz_vector_body(st, v, m, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ __launch_bounds__(1024) void lanczos_bound_kernel(LanczosState* st, LanczosVecs v, int m,
                                                             double margin) {
  bound_body(st, v, m, margin);
}

// ------------------------------------------------------------------ batched
__global__ __launch_bounds__(1024) void lanczos_init_batched(const LanczosJob* jobs, uint64_t seed) {
  const LanczosJob& J = jobs[blockIdx.x];
  init_body(J.st, J.v, J.n, J.m, seed);
}
__global__ __launch_bounds__(1024) void lanczos_step_batched(const LanczosJob* jobs, int k, double tol) {
  const LanczosJob& J = jobs[blockIdx.x];
  step_body(J.st, J.v, k, J.m, J.max_iter, tol);
}
__global__ __launch_bounds__(64) void lanczos_ritz_batched(const LanczosJob* jobs) { ritz_body(jobs[blockIdx.x].st); }
__global__ __launch_bounds__(256) void ritz_vector_batched(const LanczosJob* jobs) {
  const LanczosJob& J = jobs[blockIdx.y];
  ritz_vector_body(J.st, J.v, J.m, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ __launch_bounds__(1024) void lanczos_bound_batched(const LanczosJob* jobs, double margin) {
  const LanczosJob& J = jobs[blockIdx.x];
  bound_body(J.st, J.v, J.m, margin);
}

// Fixed scale (no bound): used by the non-auto-scaled entry points.
__global__ void set_scale_kernel(LanczosState* st, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->scale = scale;
    st->inv_scale = 1.0 / scale;
    st->nonfinite = 0;
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

// Enqueues the whole bound.  A: symmetric m x m (row-major view = column-major),
// logical size n <= m (entries beyond n are zero).  Vectors have length m.
hipError_t launch_lanczos(const void* A, int a_f64, int64_t ld, int n, int m, int max_iter, double tol,
                          uint64_t seed, double margin, LanczosState* st, const LanczosVecs& v,
                          hipStream_t s) {
  hipLaunchKernelGGL(lanczos_init_kernel, dim3(1), dim3(1024), 0, s, st, v, n, m, seed);
  for (int k = 0; k < max_iter; ++k) {
    if (a_f64) {
      launch_gemv((const double*)A, ld, m, v.q, v.w1, &st->done, s);
      launch_gemv((const double*)A, ld, m, v.w1, v.w, &st->done, s);
    } else {
      launch_gemv((const float*)A, ld, m, v.q, v.w1, &st->done, s);
      launch_gemv((const float*)A, ld, m, v.w1, v.w, &st->done, s);
    }
    hipLaunchKernelGGL(lanczos_step_kernel, dim3(1), dim3(1024), 0, s, st, v, k, m, max_iter, tol);
  }
  hipLaunchKernelGGL(lanczos_ritz_kernel, dim3(1), dim3(64), 0, s, st);
  hipLaunchKernelGGL(ritz_vector_kernel, dim3((m + 255) / 256), dim3(256), 0, s, st, v, m);
  if (a_f64)
    launch_gemv((const double*)A, ld, m, v.y, v.ay, &st->skip_ritz, s);
  else
    launch_gemv((const float*)A, ld, m, v.y, v.ay, &st->skip_ritz, s);
  hipLaunchKernelGGL(lanczos_bound_kernel, dim3(1), dim3(1024), 0, s, st, v, m, margin);
  return hipGetLastError();
}

// Batched bound over `count` float32 padded matrices (jobs in device memory;
// every job's m is a multiple of 4 and at most max_m).
// rows: one int2 {matrix, row group} per GEMV_ROWS_PER_BLOCK rows of every
// matrix (nrows entries).
hipError_t launch_lanczos_batched(const LanczosJob* jobs, int count, int max_m, int max_iter, double tol,
                                  uint64_t seed, double margin, const int2* rows, int nrows, hipStream_t s) {
  const dim3 ggrid(nrows);
  hipLaunchKernelGGL(lanczos_init_batched, dim3(count), dim3(1024), 0, s, jobs, seed);
  for (int k = 0; k < max_iter; ++k) {
    hipLaunchKernelGGL(gemv_batched_kernel, ggrid, dim3(256), 0, s, jobs, rows, 0);
    hipLaunchKernelGGL(gemv_batched_kernel, ggrid, dim3(256), 0, s, jobs, rows, 1);
    hipLaunchKernelGGL(lanczos_step_batched, dim3(count), dim3(1024), 0, s, jobs, k, tol);
  }
  hipLaunchKernelGGL(lanczos_ritz_batched, dim3(count), dim3(64), 0, s, jobs);
  hipLaunchKernelGGL(ritz_vector_batched, dim3((max_m + 255) / 256, count), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(gemv_batched_kernel, ggrid, dim3(256), 0, s, jobs, rows, 2);
  hipLaunchKernelGGL(lanczos_bound_batched, dim3(count), dim3(1024), 0, s, jobs, margin);
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
