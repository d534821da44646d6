// The fused layer step for callers that cannot allocate device memory: code
// under stream capture, tight serving loops, harnesses that forbid hipMalloc.
//
// spk_stdp_layer_fused_noalloc runs the same kernels as spk_stdp_layer_fused
// (the whole-sample path), but its buffers -- FP16 weight fragments,
// per-pixel summaries, winner keys, first-spike map, STDP bit records -- are
// carved from `scratch`, a file-scope device array of SCRATCH_BYTES, instead
// of the stream's workspace.  Each GPU holds its own copy of the array for the
// life of the process; its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a stand-in).
// Every call on a GPU shares that GPU's array and leaves its contents behind;
// every buffer carved from it is fully written by the step before it is read,
// so nothing needs clearing.  Calls that overlap on one GPU (two streams)
// race on it.
#include <mutex>
#include <vector>

#include "fused.hpp"

namespace spk
{
namespace
{
constexpr int64_t SCRATCH_BYTES = int64_t(16) << 20;

__device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// The device address of this device's scratch, looked up on the first call
// made with that device current.
uint8_t *device_scratch()
{
    static std::mutex mu;
    static std::vector<uint8_t *> ptrs;
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (int(ptrs.size()) <= dev) ptrs.resize(size_t(dev) + 1, nullptr);
    if (!ptrs[size_t(dev)])
    {
        void *p = nullptr;
        SPK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SPK_REQUIRE(p != nullptr, SPK_ERR_RUNTIME, "Fused layer: no device address for the scratch array.");
        ptrs[size_t(dev)] = static_cast<uint8_t *>(p);
    }
    return ptrs[size_t(dev)];
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_stdp_layer_fused_noalloc(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w,
                                     void *kernel, int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl,
                                     int64_t pb, int64_t pr, double threshold, int64_t ry, int64_t rx, int64_t count,
                                     const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled,
                                     int64_t pkh, int64_t pkw, int64_t psh, int64_t psw, spk_winner *winners,
                                     int64_t *counts, void *potentials, void *stream)
{
    return guard([&] {
        uint8_t *ptr = device_scratch();
        stdp_layer_fused(in, n, t, cin, h, w, kernel, cout, kh, kw, pt, pl, pb, pr, threshold, ry, rx, count, configs,
                         nconfigs, out, pooled, pkh, pkw, psh, psw, winners, counts, potentials, stream, ptr,
                         SCRATCH_BYTES);
    });
}
}
