// Allocation-free variant of kpc_dcov_components for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc).  Inputs and outputs are the caller's device buffers; the
// Lanczos basis slabs and the per-item offset / status tables live in one
// file-scope device array that each GPU holds until the process ends.  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// race; nothing in it needs to start at zero (the kernel writes every slab
// entry before reading it, and the offset tables are filled by a kernel on the
// caller's stream).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "kpc_common.h"
#include "kpc_device.h"
#include "kpcmi.h"

namespace kpc {
void set_last_error(const char* msg);
}

namespace {

constexpr long long kScratchBytes = 64LL << 20;  // 64 MiB per GPU
constexpr int kMaxItems = 65536;                 // components per call
// offset (8 B), slot (4 B) and status (4 B) tables, rounded to 256 B
constexpr long long kTableBytes = ((long long)kMaxItems * 16 + 255) / 256 * 256;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// Per-device address of `scratch` (hipGetSymbolAddress once per device, with
// that device current); the array's name is not a device address in host code.
constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};

void* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return nullptr;
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

__global__ void item_tables_kernel(long long* src_off, int* dst_slot, int count, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    src_off[i] = (long long)i * n;
    dst_slot[i] = i;
  }
}

int fail(int code, const char* msg) {
  kpc::set_last_error(msg);
  return code;
}

}  // namespace

extern "C" int kpc_dcov_components_noalloc(const double* x, int n, int count, int k,
                                           double* values, double* vectors, double* dsum,
                                           double* moment, void* stream) {
  if (!x || !values || !vectors || !dsum || !moment)
    return fail(-1, "kpc_dcov_components_noalloc: null device pointer");
  if (count <= 0 || n < 6 || n > kpc::kMaxN || k <= 0 || k >= n || k > kpc::kMaxNumCol)
    return fail(-1, "kpc_dcov_components_noalloc: invalid arguments");
  if (count > kMaxItems)
    return fail(-3, "kpc_dcov_components_noalloc: count exceeds the 65536 components the "
                    "static scratch table holds");
  const int mmax = kpc::lanczos_mmax(n);
  const long long stride = kpc::dcov_lanczos_scratch_doubles(n, k, mmax);
  const long long slabs = (kScratchBytes - kTableBytes) / (long long)sizeof(double) / stride;
  if (slabs < 1)
    return fail(-3, "kpc_dcov_components_noalloc: one Lanczos basis exceeds the 64 MiB scratch");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(-2, "kpc_dcov_components_noalloc: no current device");
  char* base = static_cast<char*>(scratch_ptr(dev));
  if (!base) return fail(-2, "kpc_dcov_components_noalloc: hipGetSymbolAddress failed");
  long long* src_off = reinterpret_cast<long long*>(base);
  int* dst_slot = reinterpret_cast<int*>(base + (long long)kMaxItems * 8);
  int* status = dst_slot + kMaxItems;
  double* slab = reinterpret_cast<double*>(base + kTableBytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(item_tables_kernel, dim3((count + 255) / 256), dim3(256), 0, s, src_off,
                     dst_slot, count, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-2, hipGetErrorString(e));
  const int grid = (int)(count < slabs ? count : slabs);
  e = kpc::launch_dcov_lanczos(x, src_off, dst_slot, count, n, k, values, vectors, dsum, moment,
                               status, slab, grid, mmax, 1e-13, s);
  if (e != hipSuccess) return fail(-2, hipGetErrorString(e));
  return 0;
}
