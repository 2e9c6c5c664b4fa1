// Fixed device scratch for the allocation-free entry point
// (rbg_profile_block_noalloc).
//
// The array lives in the library's code object, so every GPU the library
// runs on holds its own copy, at its own address, from the first use on that
// device until the process ends.  Host code never passes the array's name to
// a kernel or a copy (on the host it is only a stand-in for the symbol): the
// device address of each GPU's copy is looked up once per device, with that
// device current, when a context is opened there, and that address is what
// the host plan carves its buffers from.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB per device
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_setup_lock;
bool g_resolved[kMaxDevices];
hipError_t g_status[kMaxDevices];
void *g_address[kMaxDevices];

}  // namespace

extern "C" size_t rbg_scratch_capacity(void) { return kScratchBytes; }

// Device address of `device`'s copy of the scratch array; the caller has made
// `device` current.  Resolved once per device and kept (also a failure, whose
// status is returned again on every later call).
extern "C" hipError_t rbg_scratch_setup(int device, void **address) {
    *address = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup_lock);
    if (!g_resolved[device]) {
        void *ptr = nullptr;
        g_status[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_address[device] = g_status[device] == hipSuccess ? ptr : nullptr;
        g_resolved[device] = true;
    }
    *address = g_address[device];
    return g_status[device];
}
