// psd_projection.hpp -- C++ drop-ins with the reference's names and argument
// order, header-only over the C ABI of psdp.h.
//
// Reference: ComputationalRobotics/psd_projection, include/psd_projection/*.h.
// The cuBLAS and cuSOLVER handles of the reference become one psdp_handle_t
// (the overloads below take it in both handle positions, so call sites keep
// their shape); matrices are device pointers, column-major, in place.  As in
// the reference (include/psd_projection/check.h:18-57) a failing call prints
// the error with its location and returns; use the psdp_* C functions to get
// the status instead.
#ifndef PSD_PROJECTION_PSD_PROJECTION_HPP
#define PSD_PROJECTION_PSD_PROJECTION_HPP

#include <cstddef>
#include <cstdio>

#include "psd_projection/psdp.h"

namespace psd_projection {

namespace detail {
inline void report(psdp_status_t s, const char* what) {
  if (s != PSDP_OK) std::fprintf(stderr, "psdp error %s in %s: %s\n", psdp_status_string(s), what, psdp_last_error());
}
// The reference sizes workspaces as 3 n^2 floats; this library's layout
// (padded rows, operand pieces, Lanczos vectors) needs psdp_workspace_size
// bytes.  A caller buffer that large is used; a smaller one (a reference-
// sized buffer always is) is not, and the call takes the handle's arena
// instead, which allocates device memory on first use: that is said once
// per process on stderr, so a caller who sized the buffer to avoid
// allocation learns that it did not.
inline size_t usable_ws(int method, int64_t n, const void* ws, size_t ws_bytes) {
  if (!ws) return 0;
  psdp_config_t cfg;
  psdp_default_config(&cfg, method);
  size_t need = 0;
  if (psdp_workspace_size(&cfg, n, &need) != PSDP_OK) return 0;
  if (ws_bytes >= need) return ws_bytes;
  static bool told = false;
  if (!told) {
    told = true;
    std::fprintf(stderr,
                 "psd_projection: workspace of %zu bytes is below the %zu bytes this library needs at n = %lld "
                 "(psdp_workspace_size); using the handle's arena instead\n",
                 ws_bytes, need, (long long)n);
  }
  return 0;
}
}  // namespace detail

// composite_FP32.h:12-17
inline void composite_FP32(psdp_handle_t h, double* mat, const int mat_size, float* workspace = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP32, mat_size, workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp32(h, mat, mat_size, b ? workspace : nullptr, b), "composite_FP32");
}

// composite_FP32.h:24-30
inline void composite_FP32_auto_scale(psdp_handle_t h, psdp_handle_t /*solverH*/, double* mat, const int mat_size,
                                      float* workspace = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP32, mat_size, workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp32_auto_scale(h, mat, mat_size, b ? workspace : nullptr, b),
                 "composite_FP32_auto_scale");
}

// composite_FP16.h:12-18 (the half workspace of the reference is not needed)
inline void composite_FP16(psdp_handle_t h, double* mat, const int n, float* float_workspace = nullptr,
                           void* /*half_workspace*/ = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP16, n, float_workspace,
                                     3 * sizeof(float) * (size_t)n * n);
  detail::report(psdp_composite_fp16(h, mat, n, b ? float_workspace : nullptr, b), "composite_FP16");
}

// composite_FP16.h:26-33
inline void composite_FP16_auto_scale(psdp_handle_t h, psdp_handle_t /*solverH*/, double* mat, const int mat_size,
                                      float* float_workspace = nullptr, void* /*half_workspace*/ = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP16, mat_size, float_workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp16_auto_scale(h, mat, mat_size, b ? float_workspace : nullptr, b),
                 "composite_FP16_auto_scale");
}

// composite_FP32_emulated.h:17-22 (no CUDA-version guard: always available)
inline void composite_FP32_emulated(psdp_handle_t h, double* mat, const int mat_size, float* workspace = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP32_EMULATED, mat_size, workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp32_emulated(h, mat, mat_size, b ? workspace : nullptr, b),
                 "composite_FP32_emulated");
}

// composite_FP32_emulated.h:29-35
inline void composite_FP32_emulated_auto_scale(psdp_handle_t h, psdp_handle_t /*solverH*/, double* mat,
                                               const int mat_size, float* workspace = nullptr) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP32_EMULATED, mat_size, workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp32_emulated_auto_scale(h, mat, mat_size, b ? workspace : nullptr, b),
                 "composite_FP32_emulated_auto_scale");
}

// lanczos.h:11-17
inline void approximate_two_norm(psdp_handle_t h, psdp_handle_t /*cusolverH*/, const double* A, size_t n,
                                 double* lo, double* up, size_t max_iter = 20, double tol = 1e-10) {
  detail::report(psdp_approximate_two_norm(h, A, (int64_t)n, lo, up, (int64_t)max_iter, tol),
                 "approximate_two_norm");
}

// eig_FP64_psd.h:14-20.  As in the reference, with return_eigenvalues the
// eigenvalues (ascending, before clamping) come back in a device array of n
// doubles that the caller frees (hipFree or psdp_free); without it the
// result is nullptr (the reference returns an uninitialised pointer there,
// src/eig_FP64_psd.cu:30,84).
inline double* eig_FP64_psd(psdp_handle_t h, psdp_handle_t /*cublasH*/, double* mat, size_t mat_size,
                            bool return_eigenvalues = false) {
  double* w = nullptr;
  detail::report(psdp_eig_fp64_psd_alloc(h, mat, (int64_t)mat_size, return_eigenvalues ? &w : nullptr),
                 "eig_FP64_psd");
  return w;
}

// eig_FP32_psd.h:14-20
inline double* eig_FP32_psd(psdp_handle_t h, psdp_handle_t /*cublasH*/, double* mat, size_t mat_size,
                            bool return_eigenvalues = false) {
  double* w = nullptr;
  detail::report(psdp_eig_fp32_psd_alloc(h, mat, (int64_t)mat_size, return_eigenvalues ? &w : nullptr),
                 "eig_FP32_psd");
  return w;
}

// composite_FP32.h:36-45.  verbose: diagnostics come back through the
// status and psdp_last_error() instead of stdout.
inline void composite_FP32_auto_scale_deflate(psdp_handle_t h, psdp_handle_t /*solverH*/, double* mat,
                                              const int mat_size, const size_t k = 40, const double tol = 1e-10,
                                              const double ortho_tol = 1e-3, const bool /*verbose*/ = false) {
  detail::report(psdp_composite_fp32_auto_scale_deflate(h, mat, mat_size, (int64_t)k, tol, ortho_tol),
                 "composite_FP32_auto_scale_deflate");
}

// composite_FP16.h:39-50
inline void composite_FP16_auto_scale_deflate(psdp_handle_t h, psdp_handle_t /*solverH*/, double* mat,
                                              const int mat_size, float* float_workspace = nullptr,
                                              void* /*half_workspace*/ = nullptr, const size_t k = 40,
                                              const int maxiter = 100, const double tol = 1e-10,
                                              const bool /*verbose*/ = false) {
  const size_t b = detail::usable_ws(PSDP_METHOD_COMPOSITE_FP16, mat_size, float_workspace,
                                     3 * sizeof(float) * (size_t)mat_size * mat_size);
  detail::report(psdp_composite_fp16_auto_scale_deflate(h, mat, mat_size, b ? float_workspace : nullptr, b,
                                                        (int64_t)k, maxiter, tol),
                 "composite_FP16_auto_scale_deflate");
}

// eig_FP64_psd.h:22-31
inline void eig_FP64_deflate(psdp_handle_t h, psdp_handle_t /*cublasH*/, double* mat, size_t mat_size,
                             const size_t k = 40, const int maxiter = 100, const double tol = 1e-10,
                             const bool /*verbose*/ = false) {
  detail::report(psdp_eig_fp64_deflate(h, mat, (int64_t)mat_size, (int64_t)k, maxiter, tol), "eig_FP64_deflate");
}

// lanczos.h:33-43.  Returns the bound, as the reference does; *r receives the
// number of pairs kept (eigenvalues: k device values, eigenvectors: n x k).
inline double compute_eigenpairs(psdp_handle_t h, psdp_handle_t /*cusolverH*/, const double* A, size_t n,
                                 const size_t k, size_t* r, double* eigenvalues, double* eigenvectors,
                                 const bool upper_bound_only = false, const size_t max_iter = 0,
                                 const double tol = 1e-10, const double ortho_tol = 1e-5,
                                 const bool /*verbose*/ = false) {
  int64_t rr = 0;
  double bound = 0.0;
  detail::report(psdp_compute_eigenpairs(h, A, (int64_t)n, (int64_t)k, &rr, eigenvalues, eigenvectors,
                                         upper_bound_only ? 1 : 0, (int64_t)max_iter, tol, ortho_tol, &bound),
                 "compute_eigenpairs");
  if (r && !upper_bound_only) *r = (size_t)rr;
  return bound;
}

// lobpcg.h:16-28 (D receives m values)
inline void lobpcg(psdp_handle_t h, psdp_handle_t /*cusolverH*/, const double* A, double* V, double* D, const int n,
                   const int m, const bool warmstart = false, const int maxiter = 100, const double tol = 1e-8,
                   const bool /*verbose*/ = false) {
  detail::report(psdp_lobpcg(h, A, V, D, n, m, warmstart ? 1 : 0, maxiter, tol), "lobpcg");
}

}  // namespace psd_projection

#endif  // PSD_PROJECTION_PSD_PROJECTION_HPP
