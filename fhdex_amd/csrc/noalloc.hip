// noalloc.hip -- the device scratch of fhdx_rk3_step_noalloc.
//
// A file-scope device array, one copy per GPU the library runs on, held until the process ends.
// Host code never uses the array's name as an address (on the host it is only a stand-in
// symbol): fhdx_ctx_create looks up each device's copy once with hipGetSymbolAddress, with that
// device current, and every later use goes through the pointer kept here.
#include <mutex>

#include "fused.hpp"

namespace fhdx {

static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the no-allocation scratch table");
        return FHDX_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return FHDX_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") + hipGetErrorString(e));
        return FHDX_ERR_HIP;
    }
    g_ptr[device] = (double*)ptr;
    return FHDX_OK;
}

double* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace fhdx
