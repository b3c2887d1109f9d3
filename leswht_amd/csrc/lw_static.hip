// The static device area of the no-allocation entry point (lw_create_static).
//
// `scratch` is a file-scope device array: each GPU that loads this code
// object holds its own copy, at its own address, until the process ends. In
// host code the name `scratch` is only a host-side stand-in, so the area's
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to kernels, copies and
// fills. One handle at a time may live in a device's area: every handle
// created there would share the same bytes.
#include <hip/hip_runtime.h>

#include <mutex>

namespace lw {

constexpr size_t kStaticBytes = (size_t)256 << 20;
constexpr int kStaticMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

static std::mutex g_static_mu;
static void *g_static_ptr[kStaticMaxDevices] = {};
static bool g_static_live[kStaticMaxDevices] = {};

size_t static_area_bytes() { return kStaticBytes; }

// Claims the current device's area for one handle. Returns hipSuccess with
// the device address and size, hipErrorInvalidDevice for a device index
// beyond the table, or hipErrorAlreadyAcquired while another handle holds it.
hipError_t static_area_acquire(int *dev_out, void **ptr, size_t *bytes) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kStaticMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (g_static_live[dev]) return hipErrorAlreadyAcquired;
  if (!g_static_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_static_ptr[dev] = p;
  }
  g_static_live[dev] = true;
  *dev_out = dev;
  *ptr = g_static_ptr[dev];
  *bytes = kStaticBytes;
  return hipSuccess;
}

void static_area_release(int dev) {
  if (dev < 0 || dev >= kStaticMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_static_mu);
  g_static_live[dev] = false;
}

}  // namespace lw
