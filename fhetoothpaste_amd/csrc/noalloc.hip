// tp_bootstrap_batch_device_noalloc: the full bootstrap of tp_bootstrap_batch_device with its one
// scratch buffer (the extracted LWEs between the blind rotation and the key switch, [count][N + 1]
// words) in a file-scope device array instead of a context allocation, so that a call allocates
// nothing.  For callers that cannot call hipMalloc (code under HIP graph capture, allocation-free
// serving loops).
//
// Every GPU the library runs on holds one copy of the array (kScratchWords words, 36 MiB) from the
// load of this code object until the process ends.  Its device address on each device is resolved
// once, with that device current, when the first context on it is created (tp_context_create
// calls noalloc_setup); host code then uses only that address, never the array's name (which on the
// host is a stand-in for the symbol, not a device address).
//
// Contents: the blind rotation writes all N + 1 words of every row it uses before the key switch
// reads them, so nothing has to start at zero.  All calls on one GPU share the array: calls that
// overlap on one GPU (other host threads, other streams) race and must not be issued.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/toothpaste_amd.h"
#include "context_internal.hpp"

namespace {

constexpr size_t kScratchWords = (size_t)36 << 17;  // 36 MiB of uint64_t
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

std::mutex g_mu;
uint64_t* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_tried[kMaxDevices] = {};

uint64_t* scratch_for(int device, hipError_t& err) {
    std::lock_guard<std::mutex> lk(g_mu);
    err = g_tried[device] ? g_err[device] : hipErrorNotInitialized;
    return g_ptr[device];
}

}  // namespace

namespace tp {

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(TP_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return TP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_tried[device] = true;
    g_err[device] = e;
    if (e != hipSuccess || !p) {
        return set_error(TP_E_DEVICE,
                         (std::string("no-allocation scratch: hipGetSymbolAddress: ") + hipGetErrorString(e)).c_str());
    }
    g_ptr[device] = (uint64_t*)p;
    return TP_OK;
}

}  // namespace tp

extern "C" int tp_bootstrap_batch_device_noalloc(tp_context* c, size_t count, const uint64_t* lwes,
                                                 const uint64_t* tv, size_t tv_stride, uint64_t* out) {
    if (!c || !lwes || !tv || !out) return tp::set_error(TP_E_INVALID, "null argument");
    tp_params p{};
    tp::context_params(c, &p);
    const size_t row = (size_t)p.ring_dim + 1;
    if (count > kScratchWords / row)
        return tp::set_error(TP_E_INVALID,
                             (std::string("count exceeds the no-allocation scratch: at most ") +
                              std::to_string(kScratchWords / row) + " ciphertexts at this ring dimension")
                                 .c_str());
    const int device = tp::context_device(c);
    if (device < 0 || device >= kMaxDevices) return tp::set_error(TP_E_DEVICE, "device index out of range");
    int rc = tp::bind_device(c);
    if (rc) return rc;
    hipError_t err;
    uint64_t* ext = scratch_for(device, err);
    if (!ext)
        return tp::set_error(TP_E_DEVICE, (std::string("no-allocation scratch not set up on this device: ") +
                                           hipGetErrorString(err))
                                              .c_str());
    return tp::bootstrap_into(c, count, lwes, tv, tv_stride, ext, out);
}
