// Extremal eigenpairs and rank-r updates for the deflated projections.
//
// lanczos_pairs: compute_eigenpairs (src/lanczos.cu:319-547, declared at
// include/psd_projection/lanczos.h:33-43).
//   * Lanczos on A with the whole basis Q = [q_0 .. q_max_iter] kept and no
//     reorthogonalisation: w = A q_m; w -= beta_{m-1} q_{m-1};
//     alpha_m = q_m . w; w -= alpha_m q_m; beta_m = ||w||; stop when
//     beta_m <= DBL_EPSILON; q_{m+1} = w (1/beta_m) (src/lanczos.cu:359-389).
//     Per iteration one GEMV and one single-workgroup vector step, every
//     scalar in device memory; the host polls a flag copied to pinned memory
//     every POLL iterations and stops enqueuing once the recurrence broke
//     (launches already queued after the break are no-ops).
//   * eigenpairs of T = tridiag(alpha, beta) with rocSOLVER's tridiagonal
//     divide and conquer (dstedc); the reference runs dsyevd on the dense T
//     (src/lanczos.cu:391-423).  Eigenvalues ascending.
//   * Ritz residuals res_j = |beta_last Z(row, j)| with row 0, as the
//     reference reads Z[j*m] (src/lanczos.cu:295-301), or row nb-1, the
//     Lanczos residual ||A x_j - theta_j x_j|| of an orthonormal basis.
//   * pairs in order of decreasing |theta| (src/lanczos.cu:438-447); a pair is
//     kept when res_j < ortho_tol and |x_a . x_j| <= ortho_tol for every kept
//     x_a, x_j = Q z_j unnormalised, until k are kept (:453-507).  The
//     candidates are formed a chunk at a time, Xc = Q Zc with one GEMM, their
//     overlaps [X_kept Xc]^T Xc with two more, and the greedy test runs on the
//     host over those overlaps: the numbers the reference gets from one GEMV
//     per candidate.
//   * bound = |theta_max| + res_max + 1e-3 for the pair of largest |theta|
//     (src/lanczos.cu:511-547).
//   * normalize: the kept vectors are divided by their norms (the diagonal of
//     the candidates' Gram matrix) after the selection, which is unchanged.
//     Once the basis has lost orthogonality Q z is not a unit vector, and
//     X - theta x x^T with the reference's unnormalised x removes
//     theta ||x||^2 instead of theta.
//
// lobpcg_run: lobpcg (src/lobpcg.cu:12-256, include/psd_projection/lobpcg.h:16-28)
// for the m largest eigenpairs of sign*A.  Per iteration:
//   R = AX - X diag(lam), stop when ||R||_F < tol;
//   S = [X, R, dX]; Householder QR (rocSOLVER dgeqrf + dorgqr) -> Q;
//   AQ = sign A Q (the one n x n x 3m GEMM); T = Q^T AQ, T = (T + T^T)/2;
//   dsyevd(T); the m largest Ritz pairs in descending order:
//   X' = Q Z_m, AX' = AQ Z_m, dX = X' - X.
// Differences from the reference, none of which changes what it converges
// to: AX' is carried from AQ instead of recomputed as A X' (the reference
// spends a second n x n x m GEMM per iteration on it); the first iteration
// uses S = [X, R] (the reference stacks [X, R, X], whose third block is
// linearly dependent and only adds rounding-noise directions); and the
// initial Rayleigh-Ritz pairs each Ritz vector with its own Ritz value (the
// reference orders the vectors ascending and the values descending there,
// src/lobpcg.cu:124-133, which only inflates the first residual).  sign = -1
// gives the smallest eigenpairs of A as the largest of -A without negating
// the matrix in memory (src/composite_FP16.cu:316-319 negates it twice).
//
// rank_update: X += V diag(c) V^T, the Dger loops of the deflated projections
// (src/composite_FP32.cu:216-253, src/composite_FP16.cu:306-370,
// src/eig_FP64_psd.cu:113-170), as one column scaling and one GEMM.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <vector>

#include "psdp_device.h"
#include "psdp_eigen.h"
#include "psdp_host.h"

namespace psdp {

hipError_t launch_gemv_f64(const double* A, int64_t ld, int m, const double* x, double* y, const int* skip,
                           hipStream_t s);

struct LzFlags {
  int nb;         // completed iterations
  int done;       // recurrence stopped
  int nonfinite;  // non-finite alpha or beta met
  int pad;
};

// ------------------------------------------------------------------ kernels
// q_0 = u / ||u||, u_i ~ U(0,1] (src/lanczos.cu:344-355; a seeded stream
// stands in for cuRAND), into column 0 of Q.
__global__ __launch_bounds__(1024) void lz_init_kernel(double* __restrict__ Q, int n, uint64_t seed,
                                                       LzFlags* fl) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double q = uniform01(seed, (uint64_t)i);
    Q[i] = q;
    s += q * q;
  }
  const double nrm = sqrt(block_sum(s, red));
  if (nrm != 0.0) {
    const double inv = 1.0 / nrm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) Q[i] *= inv;
  }
  if (threadIdx.x == 0) {
    fl->nb = 0;
    fl->done = 0;
    fl->nonfinite = 0;
  }
}

// Iteration m after w = A q_m (src/lanczos.cu:366-388).
__global__ __launch_bounds__(1024) void lz_step_kernel(double* __restrict__ Q, double* __restrict__ w, int n,
                                                       int m, double* __restrict__ alpha,
                                                       double* __restrict__ beta, LzFlags* fl) {
  __shared__ double red[16];
  if (fl->done) return;
  const double* q = Q + (int64_t)m * n;
  const double* qp = Q + (int64_t)(m > 0 ? m - 1 : 0) * n;
  const double mb = m > 0 ? -beta[m - 1] : 0.0;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double wi = w[i];
    if (m > 0) wi = fma(mb, qp[i], wi);
    w[i] = wi;
    s += q[i] * wi;
  }
  const double a = block_sum(s, red);
  s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double wi = fma(-a, q[i], w[i]);
    w[i] = wi;
    s += wi * wi;
  }
  const double b = sqrt(block_sum(s, red));
  const bool bad = !isfinite(a) || !isfinite(b);
  const bool stop = bad || b <= DBL_EPSILON;
  if (!stop) {
    const double inv = 1.0 / b;
    double* qn = Q + (int64_t)(m + 1) * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) qn[i] = w[i] * inv;
  }
  if (threadIdx.x == 0) {
    alpha[m] = a;
    beta[m] = b;
    fl->nb = m + 1;
    if (stop) fl->done = 1;
    if (bad) fl->nonfinite = 1;
  }
}

// out[j] = C(row, j), j < cols (C column-major, leading dimension ldc).
__global__ void gather_row_kernel(const double* __restrict__ C, int ldc, int cols, int row,
                                  double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < cols) out[j] = C[(int64_t)j * ldc + row];
}

// out(:, c) = C(:, idx[c]), c < cnt (rows x cnt, column-major).
__global__ void gather_cols_kernel(const double* __restrict__ C, int ldc, int rows, const int* __restrict__ idx,
                                   int cnt, double* __restrict__ out) {
  const int64_t total = (int64_t)rows * cnt;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int c = (int)(e / rows);
    const int i = (int)(e - (int64_t)c * rows);
    out[e] = C[(int64_t)idx[c] * ldc + i];
  }
}

__global__ void fill_uniform_kernel(double* __restrict__ X, int64_t count, uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride)
    X[e] = uniform01(seed, (uint64_t)e);
}

// R = AX - X diag(lam) (src/lobpcg.cu:137-155: the product X lam is rounded
// before the subtraction, as Ddgmm + Daxpy do) and per-block sums of R^2.
__global__ __launch_bounds__(256) void lobpcg_residual_kernel(const double* __restrict__ AX,
                                                              const double* __restrict__ X,
                                                              const double* __restrict__ lam,
                                                              double* __restrict__ R, int n, int m,
                                                              double* __restrict__ partial) {
  __shared__ double red[4];
  const int64_t total = (int64_t)n * m;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int col = (int)(e / n);
    const double xt = X[e] * lam[col];
    const double r = AX[e] - xt;
    R[e] = r;
    s += r * r;
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

struct LobpcgScalars {
  double rnorm;
  int info;
  int pad;
};

__global__ __launch_bounds__(1024) void lobpcg_norm_kernel(const double* __restrict__ partial, int count,
                                                           const int* __restrict__ info, LobpcgScalars* out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) {
    out->rnorm = sqrt(t);
    out->info = *info;
  }
}

// T = (T + T^T) / 2 in place (p x p; src/lobpcg.cu:180-185 uses Dgeam with 1/2).
__global__ void symmetrize_small_kernel(double* __restrict__ T, int p) {
  const int64_t total = (int64_t)p * p;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e % p), j = (int)(e / p);
    if (i < j) {
      const double v = 0.5 * T[i + (int64_t)j * p] + 0.5 * T[j + (int64_t)i * p];
      T[i + (int64_t)j * p] = v;
      T[j + (int64_t)i * p] = v;
    }
  }
}

// Zm(:, j) = Z(:, p-1-j) and lam[j] = theta[p-1-j] for j < m: the m largest
// Ritz pairs in descending order (src/lobpcg.cu:193-218).
__global__ void top_columns_kernel(const double* __restrict__ Z, const double* __restrict__ theta, int p, int m,
                                   double* __restrict__ Zm, double* __restrict__ lam) {
  const int64_t total = (int64_t)p * m;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int j = (int)(e / p), i = (int)(e % p);
    Zm[e] = Z[(int64_t)(p - 1 - j) * p + i];
    if (i == 0) lam[j] = theta[p - 1 - j];
  }
}

// S(:, 0:m) = X', S(:, 2m:3m) = X' - X (src/lobpcg.cu:204-215).
__global__ void lobpcg_update_kernel(const double* __restrict__ Xn, const double* __restrict__ Xo,
                                     double* __restrict__ S, int64_t nm) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nm; e += stride) {
    const double xn = Xn[e];
    S[e] = xn;
    S[2 * nm + e] = xn - Xo[e];
  }
}

// W(:, i) = c_i V(:, i) with c_i = -s_i lam_i (mode REMOVE) or
// max(s_i lam_i, 0) (mode ADD_POSITIVE), s_i = +1 for i < split, -1 after.
__global__ void coef_scale_kernel(const double* __restrict__ V, const double* __restrict__ lam, int n, int r,
                                  int split, int mode, double* __restrict__ W) {
  const int64_t total = (int64_t)n * r;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int col = (int)(e / n);
    const double l = col < split ? lam[col] : -lam[col];
    const double c = mode == RANK_REMOVE ? -l : (l > 0.0 ? l : 0.0);
    W[e] = c * V[e];
  }
}

static unsigned grid_for(int64_t count, int block = 256, int64_t cap = 4096) {
  int64_t g = (count + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------ Lanczos pairs
static constexpr int POLL = 32;

static int pair_chunk(int64_t k, int64_t max_iter) {
  int64_t c = std::max<int64_t>(2 * k, 16);
  c = std::min<int64_t>(c, 256);
  return (int)std::max<int64_t>(1, std::min<int64_t>(c, max_iter));
}

size_t lanczos_pairs_ws_bytes(int64_t n, int64_t k, int64_t max_iter) {
  const size_t mi = (size_t)max_iter, c = (size_t)pair_chunk(k, max_iter);
  size_t b = 0;
  b += al(8 * (size_t)n * (mi + 1));  // Q
  b += al(8 * (size_t)n);             // w
  b += 4 * al(8 * mi);                // alpha, beta, D, E
  b += al(8 * mi);                    // row of Z
  b += al(8 * mi * mi);               // Z
  b += al(sizeof(LzFlags)) + al(sizeof(int));
  b += al(4 * c);                     // candidate indices
  b += al(8 * mi * c);                // Zc
  b += al(8 * (size_t)n * c);         // Xc
  b += al(8 * (size_t)k * c) + al(8 * c * c);  // overlaps
  return b;
}

psdp_status_t lanczos_pairs(const EigEnv& env, const double* A, int64_t n64, int64_t k, int64_t max_iter,
                            double ortho_tol, int residual_row, bool normalize, bool upper_only, void* ws,
                            double* evals, double* evecs, int64_t* r_out, double* bound_out) {
  const int n = (int)n64, mi = (int)max_iter, chunk = pair_chunk(k, max_iter);
  hipStream_t s = env.stream;
  rocblas_handle rb = (rocblas_handle)env.blas;
  RBCHK(rocblas_set_stream(rb, s));
  char* p = (char*)ws;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += al(bytes);
    return q;
  };
  double* Q = (double*)take(8 * (size_t)n * (mi + 1));
  double* w = (double*)take(8 * (size_t)n);
  double* alpha = (double*)take(8 * (size_t)mi);
  double* beta = (double*)take(8 * (size_t)mi);
  double* D = (double*)take(8 * (size_t)mi);
  double* E = (double*)take(8 * (size_t)mi);
  double* zrow = (double*)take(8 * (size_t)mi);
  double* Z = (double*)take(8 * (size_t)mi * mi);
  LzFlags* fl = (LzFlags*)take(sizeof(LzFlags));
  int* info = (int*)take(sizeof(int));
  int* idx_dev = (int*)take(4 * (size_t)chunk);
  double* Zc = (double*)take(8 * (size_t)mi * chunk);
  double* Xc = (double*)take(8 * (size_t)n * chunk);
  double* G1 = (double*)take(8 * (size_t)std::max<int64_t>(k, 1) * chunk);
  double* G2 = (double*)take(8 * (size_t)chunk * chunk);

  // ---- recurrence
  LzFlags* hf = (LzFlags*)env.pinned;  // two poll slots
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 2; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  hipLaunchKernelGGL(lz_init_kernel, dim3(1), dim3(1024), 0, s, Q, n, env.seed, fl);
  int polls = 0;
  for (int m = 0; m < mi; ++m) {
    HIPCHK(launch_gemv_f64(A, n, n, Q + (int64_t)m * n, w, &fl->done, s));
    hipLaunchKernelGGL(lz_step_kernel, dim3(1), dim3(1024), 0, s, Q, w, n, m, alpha, beta, fl);
    if ((m + 1) % POLL == 0 && m + 1 < mi) {
      // look at the poll before this one (already finished or nearly so)
      if (polls > 0) {
        const int prev = (polls - 1) & 1;
        HIPCHK(hipEventSynchronize(ev[prev]));
        if (hf[prev].done) break;
      }
      const int slot = polls & 1;
      HIPCHK(hipMemcpyAsync(&hf[slot], fl, sizeof(LzFlags), hipMemcpyDeviceToHost, s));
      HIPCHK(hipEventRecord(ev[slot], s));
      ++polls;
    }
  }
  HIPCHK(hipGetLastError());
  LzFlags f;
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(&f, fl, sizeof f, hipMemcpyDeviceToHost));
  if (f.nonfinite) return fail(PSDP_ERR_NONFINITE, "non-finite value in the Lanczos recurrence");
  const int nb = f.nb;
  if (nb < 1) return fail(PSDP_ERR_DEVICE, "Lanczos recurrence made no iteration");
  double beta_last = 0.0;
  HIPCHK(hipMemcpy(&beta_last, beta + (nb - 1), sizeof(double), hipMemcpyDeviceToHost));

  // ---- T = tridiag(alpha, beta): eigenpairs (ascending)
  HIPCHK(hipMemcpyAsync(D, alpha, 8 * (size_t)nb, hipMemcpyDeviceToDevice, s));
  if (nb > 1) HIPCHK(hipMemcpyAsync(E, beta, 8 * (size_t)(nb - 1), hipMemcpyDeviceToDevice, s));
  RBCHK(rocsolver_dstedc(rb, rocblas_evect_tridiagonal, nb, D, E, Z, nb, info));
  const int row = residual_row == PSDP_RITZ_RESIDUAL_LAST_ROW ? nb - 1 : 0;
  hipLaunchKernelGGL(gather_row_kernel, dim3(grid_for(nb)), dim3(256), 0, s, Z, nb, nb, row, zrow);
  HIPCHK(hipGetLastError());
  std::vector<double> theta(nb), zr(nb);
  int info_h = 0;
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(theta.data(), D, 8 * (size_t)nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(zr.data(), zrow, 8 * (size_t)nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&info_h, info, sizeof info_h, hipMemcpyDeviceToHost));
  if (info_h != 0) return fail(PSDP_ERR_NONFINITE, "tridiagonal eigensolver did not converge (info = %d)", info_h);
  std::vector<double> res(nb);
  for (int j = 0; j < nb; ++j) res[j] = std::fabs(beta_last * zr[j]);
  std::vector<int> idx(nb);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return std::fabs(theta[a]) > std::fabs(theta[b]); });
  *bound_out = std::fabs(theta[idx[0]]) + res[idx[0]] + 1e-3;
  if (upper_only) return PSDP_OK;

  // ---- selection (residual filter, then overlap filter)
  std::vector<int> cand;
  for (int j : idx)
    if (res[j] < ortho_tol) cand.push_back(j);
  std::vector<double> sel;
  std::vector<double> g1, g2;
  for (size_t pos = 0; pos < cand.size() && (int64_t)sel.size() < k; pos += chunk) {
    const int cnt = (int)std::min<size_t>(chunk, cand.size() - pos);
    const int acc = (int)sel.size();
    HIPCHK(hipMemcpy(idx_dev, cand.data() + pos, 4 * (size_t)cnt, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_for((int64_t)nb * cnt)), dim3(256), 0, s, Z, nb, nb,
                       idx_dev, cnt, Zc);
    HIPCHK(hipGetLastError());
    const double one = 1.0, zero = 0.0;
    RBCHK(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_none, n, cnt, nb, &one, Q, n, Zc, nb, &zero,
                        Xc, n));
    if (acc > 0)
      RBCHK(rocblas_dgemm(rb, rocblas_operation_transpose, rocblas_operation_none, acc, cnt, n, &one, evecs, n,
                          Xc, n, &zero, G1, acc));
    RBCHK(rocblas_dgemm(rb, rocblas_operation_transpose, rocblas_operation_none, cnt, cnt, n, &one, Xc, n, Xc, n,
                        &zero, G2, cnt));
    g1.resize((size_t)acc * cnt);
    g2.resize((size_t)cnt * cnt);
    HIPCHK(hipStreamSynchronize(s));
    if (acc > 0) HIPCHK(hipMemcpy(g1.data(), G1, 8 * g1.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(g2.data(), G2, 8 * g2.size(), hipMemcpyDeviceToHost));
    std::vector<int> mine;  // accepted columns of this chunk
    for (int c = 0; c < cnt && (int64_t)sel.size() < k; ++c) {
      bool ok = true;
      for (int a = 0; a < acc && ok; ++a) ok = std::fabs(g1[(size_t)c * acc + a]) <= ortho_tol;
      for (size_t t = 0; t < mine.size() && ok; ++t) ok = std::fabs(g2[(size_t)c * cnt + mine[t]]) <= ortho_tol;
      if (!ok) continue;
      double* dst = evecs + (int64_t)sel.size() * n;
      HIPCHK(hipMemcpyAsync(dst, Xc + (int64_t)c * n, 8 * (size_t)n, hipMemcpyDeviceToDevice, s));
      const double nrm2 = g2[(size_t)c * cnt + c];
      if (normalize && nrm2 > 0.0) {
        const double inv = 1.0 / std::sqrt(nrm2);
        RBCHK(rocblas_dscal(rb, n, &inv, dst, 1));
      }
      sel.push_back(theta[cand[pos + c]]);
      mine.push_back(c);
    }
  }
  if (!sel.empty()) HIPCHK(hipMemcpyAsync(evals, sel.data(), 8 * sel.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  *r_out = (int64_t)sel.size();
  return PSDP_OK;
}

// ------------------------------------------------------------------ LOBPCG
static constexpr int RES_BLOCKS = 512;

size_t lobpcg_ws_bytes(int64_t n, int64_t m) {
  const size_t nm = (size_t)n * m, p = 3 * (size_t)m;
  size_t b = 0;
  b += 4 * al(8 * nm);                  // X, X', AX, AX'
  b += 2 * al(8 * nm * 3);              // S (Q), AS
  b += al(8 * p * p);                   // T / Z
  b += 3 * al(8 * p);                   // theta, E, tau
  b += al(8 * p * m) + al(8 * (size_t)m);  // Zm, lam
  b += al(8 * RES_BLOCKS) + al(sizeof(LobpcgScalars)) + al(sizeof(int));
  return b;
}

psdp_status_t lobpcg_run(const EigEnv& env, const double* A, double sign, double* V, double* Dout, int64_t n64,
                         int64_t m64, bool warm, int maxiter, double tol, void* ws, int* iters_out,
                         double* rnorm_out) {
  const int n = (int)n64, m = (int)m64;
  const int64_t nm = n64 * m64;
  hipStream_t s = env.stream;
  rocblas_handle rb = (rocblas_handle)env.blas;
  RBCHK(rocblas_set_stream(rb, s));
  char* p = (char*)ws;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += al(bytes);
    return q;
  };
  double* X = (double*)take(8 * nm);
  double* Xn = (double*)take(8 * nm);
  double* AX = (double*)take(8 * nm);
  double* AXn = (double*)take(8 * nm);
  double* S = (double*)take(8 * nm * 3);
  double* AS = (double*)take(8 * nm * 3);
  double* T = (double*)take(8 * (size_t)9 * m * m);
  double* theta = (double*)take(8 * (size_t)3 * m);
  double* E = (double*)take(8 * (size_t)3 * m);
  double* tau = (double*)take(8 * (size_t)3 * m);
  double* Zm = (double*)take(8 * (size_t)3 * m * m);
  double* lam = (double*)take(8 * (size_t)m);
  double* partial = (double*)take(8 * RES_BLOCKS);
  LobpcgScalars* sc = (LobpcgScalars*)take(sizeof(LobpcgScalars));
  int* info = (int*)take(sizeof(int));
  const double one = 1.0, zero = 0.0;
  LobpcgScalars* hs = (LobpcgScalars*)env.pinned;

  HIPCHK(hipMemsetAsync(info, 0, sizeof(int), s));
  // Rayleigh-Ritz on the orthonormal columns Qb (n x q): T = Qb^T (sign A Qb),
  // the m largest pairs -> X' = Qb Zm, AX' = AQb Zm, lam.
  auto rayleigh_ritz = [&](const double* Qb, double* AQb, int q) -> psdp_status_t {
    RBCHK(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_none, n, q, n, &sign, A, n, Qb, n, &zero,
                        AQb, n));
    RBCHK(rocblas_dgemm(rb, rocblas_operation_transpose, rocblas_operation_none, q, q, n, &one, Qb, n, AQb, n,
                        &zero, T, q));
    hipLaunchKernelGGL(symmetrize_small_kernel, dim3(grid_for((int64_t)q * q)), dim3(256), 0, s, T, q);
    RBCHK(rocsolver_dsyevd(rb, rocblas_evect_original, rocblas_fill_upper, q, T, q, theta, E, info));
    hipLaunchKernelGGL(top_columns_kernel, dim3(grid_for((int64_t)q * m)), dim3(256), 0, s, T, theta, q, m, Zm,
                       lam);
    RBCHK(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_none, n, m, q, &one, Qb, n, Zm, q, &zero,
                        Xn, n));
    RBCHK(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_none, n, m, q, &one, AQb, n, Zm, q, &zero,
                        AXn, n));
    HIPCHK(hipGetLastError());
    return PSDP_OK;
  };

  // ---- initial block (src/lobpcg.cu:84-135)
  if (warm) {
    HIPCHK(hipMemcpyAsync(X, V, 8 * nm, hipMemcpyDeviceToDevice, s));
  } else {
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(nm)), dim3(256), 0, s, X, nm, env.seed);
    RBCHK(rocsolver_dgeqrf(rb, n, m, X, n, tau));
    RBCHK(rocsolver_dorgqr(rb, n, m, m, X, n, tau));
  }
  STCHK(rayleigh_ritz(X, AS, m));
  std::swap(X, Xn);
  std::swap(AX, AXn);
  HIPCHK(hipMemcpyAsync(S, X, 8 * nm, hipMemcpyDeviceToDevice, s));

  int iters = 0;
  double rn = 0.0;
  for (int it = 1; it <= maxiter; ++it) {
    // S(:, m:2m) = R, ||R||_F
    hipLaunchKernelGGL(lobpcg_residual_kernel, dim3(grid_for(nm, 256, RES_BLOCKS)), dim3(256), 0, s, AX, X, lam,
                       S + nm, n, m, partial);
    hipLaunchKernelGGL(lobpcg_norm_kernel, dim3(1), dim3(1024), 0, s, partial, (int)grid_for(nm, 256, RES_BLOCKS),
                       info, sc);
    HIPCHK(hipMemcpyAsync(hs, sc, sizeof(LobpcgScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    rn = hs->rnorm;
    if (hs->info != 0) return fail(PSDP_ERR_NONFINITE, "LOBPCG Rayleigh-Ritz eigensolver failed (info = %d)", hs->info);
    if (!std::isfinite(rn)) return fail(PSDP_ERR_NONFINITE, "non-finite LOBPCG residual");
    if (rn < tol) break;
    iters = it;
    const int q = it == 1 ? 2 * m : 3 * m;
    RBCHK(rocsolver_dgeqrf(rb, n, q, S, n, tau));
    RBCHK(rocsolver_dorgqr(rb, n, q, q, S, n, tau));
    STCHK(rayleigh_ritz(S, AS, q));
    hipLaunchKernelGGL(lobpcg_update_kernel, dim3(grid_for(nm)), dim3(256), 0, s, Xn, X, S, nm);
    std::swap(X, Xn);
    std::swap(AX, AXn);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(V, X, 8 * nm, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(Dout, lam, 8 * (size_t)m, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  int info_h = 0;
  HIPCHK(hipMemcpy(&info_h, info, sizeof info_h, hipMemcpyDeviceToHost));
  if (info_h != 0) return fail(PSDP_ERR_NONFINITE, "LOBPCG Rayleigh-Ritz eigensolver failed (info = %d)", info_h);
  if (iters_out) *iters_out = iters;
  if (rnorm_out) *rnorm_out = rn;
  return PSDP_OK;
}

// ------------------------------------------------------------------ rank update
psdp_status_t rank_update(const EigEnv& env, double* X, int64_t n64, const double* V, const double* lam, int r,
                          int split, int mode, double* W) {
  if (r <= 0) return PSDP_OK;
  const int n = (int)n64;
  rocblas_handle rb = (rocblas_handle)env.blas;
  RBCHK(rocblas_set_stream(rb, env.stream));
  hipLaunchKernelGGL(coef_scale_kernel, dim3(grid_for(n64 * r)), dim3(256), 0, env.stream, V, lam, n, r, split,
                     mode, W);
  HIPCHK(hipGetLastError());
  const double one = 1.0;
  RBCHK(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_transpose, n, n, r, &one, W, n, V, n, &one, X,
                      n));
  return PSDP_OK;
}

}  // namespace psdp
