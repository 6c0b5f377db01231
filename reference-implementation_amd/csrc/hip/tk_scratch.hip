// Device scratch of the allocation-free entry point
// (TAPP_AMD_execute_product_noalloc): one file-scope device array per GPU,
// allocated when the code object loads and held until the process ends.
// It holds everything the product path otherwise takes from hipMalloc: the
// plan's offset tables, pre-pass buffers, split-K partial sums and the
// split-F16 range words.
//
// The array's name is a host-side stand-in, not a device address: host code
// reaches the memory only through the address hipGetSymbolAddress returns,
// taken once per device with that device current and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../tk.h"

static __device__ __align__(256) unsigned char scratch[TK_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex mu;
void* dev_ptr[kMaxDevices];
}  // namespace

extern "C" int tk_scratch_get(int device, void** ptr, size_t* bytes)
{
    if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!dev_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        dev_ptr[device] = p;
    }
    *ptr = dev_ptr[device];
    *bytes = TK_SCRATCH_BYTES;
    return 0;
}
