// noalloc.hip — the device array behind via_answer_device_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get an answer entry
// point whose whole workspace lives in this file-scope array. Each GPU the
// library runs on holds its own copy at its own address from the library's
// first context on that device until the process ends. In host code the
// name `scratch` is a host-side stand-in, not a device address: every host
// use goes through the pointer hipGetSymbolAddress returns for the device,
// recorded once per device here.
#include <map>
#include <mutex>

#include "host_internal.h"

namespace {

static __device__ __align__(256) uint8_t scratch[via::kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> its copy of `scratch`

}  // namespace

// Called by via_ctx_create with `device` current; the address is looked up
// once per device and reused by every later context on it.
via_status via::scratch_setup(int device, void **ptr) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *ptr = it->second;
        return VIA_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return VIA_E_DEVICE;
    g_scratch_ptr[device] = p;
    *ptr = p;
    return VIA_OK;
}
