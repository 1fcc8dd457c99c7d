// Lookup without device allocation (pt_table_lookup_dev_noalloc), for callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses that forbid it.
//
// The lookup's working buffers (query descriptor, GCK planes, the two lattice fields, candidate
// history, gather output, initial field) come from `scratch`, a file-scope device array of
// PT_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process ends. In
// host code the array's name is only a stand-in for the symbol: its device address is looked up
// once per device with hipGetSymbolAddress, with that device current, and only that address is
// handed to kernels. Every call on one GPU shares that GPU's array, which keeps the last call's
// contents; the lookup writes every entry of it before reading, so nothing needs clearing. Calls
// that overlapped on one GPU would race on the array, so calls are serialised per device by a
// host lock held until the lookup has finished.
#include "pt_device.h"
#include "pt_kernels.h"

#include <mutex>

namespace {

constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[PT_NOALLOC_SCRATCH_BYTES];

std::mutex g_init_lock;
unsigned char *g_scratch_ptr[MAX_DEVICES] = {};   // device address of `scratch`, per device
std::mutex g_call_lock[MAX_DEVICES];

// Device address of this device's copy of `scratch` (looked up on first use, device current).
int scratch_address(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> g(g_init_lock);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return pt_failf(PT_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return PT_OK;
}

}  // namespace

extern "C" int pt_table_lookup_dev_noalloc(pt_table *t, const float *a_dev, int h, int w, int c,
                                           double *ann_out_dev, const double *ann_prev_dev, const float *sv_dev,
                                           double *seconds) {
    if (!t || !a_dev || !ann_out_dev) return pt_fail(PT_E_INVALID, "pt_table_lookup_dev_noalloc: null argument");
    int dev = -1;
    PT_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return pt_failf(PT_E_UNSUPPORTED, "device %d out of range", dev);
    if (pt_table_device(t) != dev)
        return pt_failf(PT_E_INVALID, "table was built on device %d, current device is %d", pt_table_device(t), dev);
    size_t need = 0;
    int rc = pt_lookup_workspace_layout(t, h, w, c, ann_prev_dev != nullptr, nullptr, &need, nullptr);
    if (rc != PT_OK) return rc;
    if (need > (size_t)PT_NOALLOC_SCRATCH_BYTES)
        return pt_failf(PT_E_INVALID,
                        "query needs %zu bytes of lookup scratch, more than the %zu of the no-allocation entry point",
                        need, (size_t)PT_NOALLOC_SCRATCH_BYTES);
    unsigned char *ws = nullptr;
    rc = scratch_address(dev, &ws);
    if (rc != PT_OK) return rc;
    std::lock_guard<std::mutex> g(g_call_lock[dev]);
    return pt_lookup_in_workspace(t, a_dev, h, w, c, ann_out_dev, ann_prev_dev, sv_dev, ws,
                                  (size_t)PT_NOALLOC_SCRATCH_BYTES, seconds);
}
