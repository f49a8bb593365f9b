// Scratch memory of the allocation-free projection (psdp_project_f32_noalloc).
//
// Some callers cannot allocate device memory inside a call: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.  Their
// entry point takes its workspace, Lanczos state and tile lists from one
// file-scope device array of fixed size, `scratch`, that every GPU the
// library runs on holds until the process ends.  Host code never uses the
// array's name as an address (on the host it is only a stand-in for the
// symbol): at handle creation the device's own address of the array is
// looked up once per device with hipGetSymbolAddress, with that device
// current, and kept per device; kernels get that address.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU race.
#include <mutex>

#include "psdp_device.h"
#include "psdp_host.h"

namespace psdp {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB per device

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};
}  // namespace

// Device address of this device's scratch array.  The caller has made
// `device` current; the lookup runs once per device.
psdp_status_t noalloc_scratch(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= MAX_DEVICES) return fail(PSDP_ERR_UNSUPPORTED, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PSDP_ERR_DEVICE, "no device address for the scratch array on device %d", device);
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  *bytes = NOALLOC_SCRATCH_BYTES;
  return PSDP_OK;
}

// Tile list of the upper triangle of a T x T tile grid in column order
// (0,0), (0,1), (1,1), (0,2), ...: entry w is tile (I, J) with
// J (J + 1) / 2 <= w < (J + 1) (J + 2) / 2, I = w - J (J + 1) / 2.
__global__ __launch_bounds__(256) void fill_tile_list_kernel(TileRef* __restrict__ out, int T, int count) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= count) return;
  int J = (int)((sqrtf(8.0f * (float)w + 1.0f) - 1.0f) * 0.5f);
  while ((J + 1) * (J + 2) / 2 <= w) ++J;
  while (J * (J + 1) / 2 > w) --J;
  out[w] = TileRef{0, w - J * (J + 1) / 2, J, 0};
}

hipError_t launch_fill_tile_list(TileRef* out, int T, hipStream_t s) {
  const int count = T * (T + 1) / 2;
  hipLaunchKernelGGL(fill_tile_list_kernel, dim3((count + 255) / 256), dim3(256), 0, s, out, T, count);
  return hipGetLastError();
}

}  // namespace psdp
