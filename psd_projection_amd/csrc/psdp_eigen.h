// Host interface of the eigenpair routines (eigen.hip) used by the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/psd_projection/psdp.h"

namespace psdp {

// What the eigenpair routines take from a handle.
struct EigEnv {
  void* blas;         // rocblas_handle
  hipStream_t stream;
  uint64_t seed;      // start vectors / start block
  void* pinned;       // >= 64 bytes of pinned host memory for polled scalars
};

enum : int { RANK_REMOVE = 0, RANK_ADD_POSITIVE = 1 };

// compute_eigenpairs (src/lanczos.cu:319-547).  ws: lanczos_pairs_ws_bytes.
// evals (k) and evecs (n x k) are device arrays; *r_out and *bound_out host.
size_t lanczos_pairs_ws_bytes(int64_t n, int64_t k, int64_t max_iter);
psdp_status_t lanczos_pairs(const EigEnv& env, const double* A, int64_t n, int64_t k, int64_t max_iter,
                            double ortho_tol, int residual_row, bool normalize, bool upper_only, void* ws,
                            double* evals, double* evecs, int64_t* r_out, double* bound_out);

// lobpcg (src/lobpcg.cu:12-256) on sign * A.  ws: lobpcg_ws_bytes.  V (n x m)
// and D (m) are device arrays.
size_t lobpcg_ws_bytes(int64_t n, int64_t m);
psdp_status_t lobpcg_run(const EigEnv& env, const double* A, double sign, double* V, double* D, int64_t n,
                         int64_t m, bool warm, int maxiter, double tol, void* ws, int* iters_out,
                         double* rnorm_out);

// X += V diag(c) V^T with c from lam (see coef_scale_kernel); W: n x r scratch.
psdp_status_t rank_update(const EigEnv& env, double* X, int64_t n, const double* V, const double* lam, int r,
                          int split, int mode, double* W);

}  // namespace psdp
