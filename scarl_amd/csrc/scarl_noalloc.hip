// scarl_noalloc.hip -- the static device scratch of the allocation-free entry point
// scarl_hip_gen_std_gauss_fft_dev_noalloc (include/scarl_hip.h).
//
// The array lives in this translation unit's code object.  Its device address is looked up once
// per device, with that device current, at the library's init call (noalloc_setup, called by
// ensure_init in scarl_runtime.hip); host code then uses only that address, never the array's
// name (which on the host is a stand-in symbol, not device memory).  Every call on one GPU shares
// the array and the last call's contents stay in it; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

namespace scarl {

// 1 MiB for the plan table and twiddle pool, 96 MiB for the two REDFT01 workspaces of a 256^3
// field ((n1 m2 m3 + n1 n2 m3) doubles = 2^22 + 2^23 doubles)
constexpr size_t kNoAllocScratchBytes = (size_t(1) << 20) + (size_t(96) << 20);
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static void* g_scratch_ptr[kNoAllocMaxDevices];
static bool g_scratch_done[kNoAllocMaxDevices];
static std::string g_scratch_err[kNoAllocMaxDevices];

// Looks up the scratch address on `device` (which must be current), once.
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    if (g_scratch_done[device]) return g_scratch_ptr[device] ? hipSuccess : hipErrorInvalidSymbol;
    g_scratch_done[device] = true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        char buf[256];
        snprintf(buf, sizeof buf, "allocation-free entry point: hipGetSymbolAddress of the scratch array failed on "
                 "device %d: %s", device, hipGetErrorString(e));
        g_scratch_err[device] = buf;
        g_scratch_ptr[device] = nullptr;
        return e != hipSuccess ? e : hipErrorInvalidSymbol;
    }
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

// Device address of the scratch on `device` (nullptr if the lookup failed) and its size.
void* noalloc_scratch(int device, size_t* bytes) {
    if (bytes) *bytes = kNoAllocScratchBytes;
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    return g_scratch_ptr[device];
}

const char* noalloc_error(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return "allocation-free entry point: device index out of range";
    if (!g_scratch_done[device]) return "allocation-free entry point: the library is not initialised on this device";
    return g_scratch_err[device].c_str();
}

}  // namespace scarl
