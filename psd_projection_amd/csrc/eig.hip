// EVD projections Pi(X) = V max(L, 0) V^T on rocSOLVER/rocBLAS: drop-ins of
// eig_FP64_psd (src/eig_FP64_psd.cu:24-85) and eig_FP32_psd
// (src/eig_FP32_psd.cu:23-96).  Not on the hot path: they are the method the
// composite chain replaces and the reference's own oracle.
//
// Sequence: syevd (jobz = V, uplo = U) on a copy of X, optional copy of the
// eigenvalues (before clamping) to the caller's buffer, column scaling
// V_s = V diag(max(w, 0)) in one kernel (the reference issues n Dscal calls),
// then X = V_s V^T with one library GEMM (N, T).  The FP32 variant rounds X to
// float first and converts the result back, as the reference does.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "psdp_device.h"

namespace psdp {

template <typename T>
__global__ void scale_columns_kernel(const T* __restrict__ V, const T* __restrict__ w, T* __restrict__ Vs,
                                     int n) {
  const int64_t count = (int64_t)n * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < count; idx += stride) {
    const int j = (int)(idx / n);
    const T s = w[j] > T(0) ? w[j] : T(0);
    Vs[idx] = V[idx] * s;
  }
}

template <typename A, typename B>
__global__ void convert_kernel(const A* __restrict__ in, B* __restrict__ out, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) out[i] = (B)in[i];
}

template <typename A, typename B>
static void convert(const A* in, B* out, int64_t count, hipStream_t s) {
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((convert_kernel<A, B>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, count);
}

// Workspace: 2 n^2 + 2 n values of T plus an int (caller sizes it with
// eig_workspace_bytes).  Returns the rocSOLVER/rocBLAS status (0 = success);
// *info_dev receives syevd's info.
template <typename T>
static int eig_psd_impl(rocblas_handle rb, hipStream_t s, T* X, int n, T* ws, int* info_dev, T* w_out) {
  const int64_t nn = (int64_t)n * n;
  T* V = ws;
  T* Vs = ws + nn;
  T* D = ws + 2 * nn;
  T* E = D + n;
  if (hipMemcpyAsync(V, X, nn * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  rocblas_status st;
  if constexpr (sizeof(T) == 8)
    st = rocsolver_dsyevd(rb, rocblas_evect_original, rocblas_fill_upper, n, V, n, D, E, info_dev);
  else
    st = rocsolver_ssyevd(rb, rocblas_evect_original, rocblas_fill_upper, n, V, n, D, E, info_dev);
  if (st != rocblas_status_success) return (int)st;
  if (w_out && hipMemcpyAsync(w_out, D, n * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  int64_t blocks = (nn + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(scale_columns_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, V, D, Vs, n);
  const T one = 1, zero = 0;
  if constexpr (sizeof(T) == 8)
    st = rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, Vs, n, V, n, &zero,
                       X, n);
  else
    st = rocblas_sgemm(rb, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, Vs, n, V, n, &zero,
                       X, n);
  return (int)st;
}

size_t eig_workspace_bytes(int n, int f64) {
  const size_t nn = (size_t)n * n;
  if (f64) return (2 * nn + 2 * (size_t)n) * 8 + 256;
  // float copy of X, V, V_s, D, E, D as double
  return (3 * nn + 2 * (size_t)n) * 4 + (size_t)n * 8 + 512;
}

// X (double, n x n column-major) <- Pi(X).  f64 = 0 computes in float like
// eig_FP32_psd.  eig (double, n) receives the eigenvalues when non-null.
int eig_project(void* rb_handle, hipStream_t s, double* X, int n, int f64, void* ws, int* info_dev,
                double* eig) {
  rocblas_handle rb = (rocblas_handle)rb_handle;
  if (rocblas_set_stream(rb, s) != rocblas_status_success) return -1;
  if (f64) return eig_psd_impl<double>(rb, s, X, n, (double*)ws, info_dev, eig);
  const int64_t nn = (int64_t)n * n;
  float* Xf = (float*)ws;
  float* rest = Xf + nn;
  float* Df = rest + 2 * nn;
  convert<double, float>(X, Xf, nn, s);
  int st = eig_psd_impl<float>(rb, s, Xf, n, rest, info_dev, nullptr);
  if (st) return st;
  if (eig) convert<float, double>(Df, eig, n, s);
  convert<float, double>(Xf, X, nn, s);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

void* eig_create_blas() {
  rocblas_handle rb = nullptr;
  if (rocblas_create_handle(&rb) != rocblas_status_success) return nullptr;
  return (void*)rb;
}

void eig_destroy_blas(void* rb) {
  if (rb) rocblas_destroy_handle((rocblas_handle)rb);
}

}  // namespace psdp
