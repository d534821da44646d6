// Runtime part of the C ABI: errors, devices and the caching allocator.
//
// The allocator follows the reference's CUDA allocator
// (src/spyker/cuda/memory.cu:40-155): blocks are cached by exact size and
// handed out again on the next request of that size; cache_clear returns
// the cached blocks to the driver.  Per-call scratch does not use it: it
// comes from the stream-ordered pool (hipMallocAsync), so a kernel still
// reading a block never sees it reused by another stream.  Entry points that
// run every training step keep a grow-only workspace per stream instead
// (stream_workspace), which costs no allocation call once warm and lets a
// stream capture record the call without memory nodes; a block a capture
// has used is kept for the life of the process.
#include <map>
#include <tuple>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace spk
{
namespace
{
thread_local std::string g_last_error;

struct Cache
{
    std::mutex mu;
    std::multimap<int64_t, void *> free;     // size -> block
    std::unordered_map<void *, int64_t> taken;  // block -> size
    int64_t cached_bytes = 0, used_bytes = 0;
};
Cache &cache()
{
    static Cache *c = new Cache;  // never destroyed: blocks may outlive static teardown
    return *c;
}
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

const char *type_name(int32_t type)
{
    static const char *names[] = {"I8", "I16", "I32", "I64", "U8", "U16", "U32", "U64", "F16", "F32", "F64"};
    return (type >= 0 && type <= 10) ? names[type] : "unknown";
}

namespace
{
// A workspace block that a stream capture has referenced is baked into the
// captured graph's kernel arguments, so it must stay valid for as long as
// the graph may be replayed: such a block is never freed (growing moves it
// to `retired`, and cache_clear keeps it) until the process ends.
struct Workspace
{
    void *ptr = nullptr;
    int64_t bytes = 0;
    bool captured = false;
};
struct Workspaces
{
    std::mutex mu;
    std::map<std::tuple<int, int, hipStream_t>, Workspace> map;  // (device, owner, stream)
    std::vector<void *> retired;                                  // captured blocks, never freed
};
Workspaces &workspaces()
{
    static Workspaces *w = new Workspaces;
    return *w;
}
}  // namespace

int cu_count()
{
    static std::mutex mu;
    static std::vector<int> cus;
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (int(cus.size()) <= dev) cus.resize(size_t(dev) + 1, 0);
    if (!cus[size_t(dev)]) SPK_HIP(hipDeviceGetAttribute(&cus[size_t(dev)], hipDeviceAttributeMultiprocessorCount, dev));
    return cus[size_t(dev)];
}

void *stream_workspace(int owner, hipStream_t stream, int64_t bytes)
{
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SPK_HIP(hipStreamIsCapturing(stream, &cap));
    Workspaces &w = workspaces();
    std::lock_guard<std::mutex> lock(w.mu);
    Workspace &e = w.map[std::make_tuple(dev, owner, stream)];
    if (e.bytes < bytes)
    {
        SPK_ARG(cap == hipStreamCaptureStatusNone,
                "The workspace of this call must grow, which a stream capture cannot do: run the call once on the "
                "capturing stream before capturing it.");
        if (e.ptr)
        {
            if (e.captured)
                w.retired.push_back(e.ptr);  // a captured graph may still read it
            else
            {
                // the old block may still be in use by work enqueued on the stream
                SPK_HIP(hipStreamSynchronize(stream));
                SPK_HIP(hipFree(e.ptr));
            }
            e = Workspace{};
        }
        SPK_HIP(hipMalloc(&e.ptr, size_t(bytes)));
        e.bytes = bytes;
    }
    if (cap != hipStreamCaptureStatusNone) e.captured = true;
    return e.ptr;
}

void clear_workspaces()
{
    Workspaces &w = workspaces();
    std::lock_guard<std::mutex> lock(w.mu);
    for (auto it = w.map.begin(); it != w.map.end();)
    {
        if (it->second.captured)
        {
            ++it;
            continue;
        }
        if (it->second.ptr) SPK_HIP(hipFree(it->second.ptr));
        it = w.map.erase(it);
    }
}

Scratch::Scratch(int64_t bytes_, hipStream_t stream_) : bytes(bytes_), stream(stream_)
{
    if (bytes <= 0) return;
    // the device's default pool keeps what is freed instead of returning it
    // to the driver at each synchronization, so steady-state calls do not
    // map memory again
    static std::mutex mu;
    static std::vector<char> done;
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lock(mu);
        if (int(done.size()) <= dev) done.resize(size_t(dev) + 1, 0);
        if (!done[size_t(dev)])
        {
            hipMemPool_t pool;
            SPK_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
            uint64_t keep = UINT64_MAX;
            SPK_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
            done[size_t(dev)] = 1;
        }
    }
    SPK_HIP(hipMallocAsync(&ptr, size_t(bytes), stream));
}

Scratch::~Scratch()
{
    if (ptr) (void)hipFreeAsync(ptr, stream);
}
}  // namespace spk

using namespace spk;

extern "C" {

const char *spk_last_error(void) { return g_last_error.c_str(); }

int32_t spk_version(void) { return 1 * 10000 + 0 * 100 + 0; }

int32_t spk_device_count(int32_t *count)
{
    return guard([&] {
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

int32_t spk_set_device(int32_t device) { return guard([&] { SPK_HIP(hipSetDevice(device)); }); }

int32_t spk_current_device(int32_t *device)
{
    return guard([&] {
        int d = 0;
        SPK_HIP(hipGetDevice(&d));
        *device = d;
    });
}

int32_t spk_memory_info(int64_t *free_bytes, int64_t *total_bytes)
{
    return guard([&] {
        size_t f = 0, t = 0;
        SPK_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = int64_t(f);
        if (total_bytes) *total_bytes = int64_t(t);
    });
}

int32_t spk_synchronize(void *stream) { return guard([&] { SPK_HIP(hipStreamSynchronize(S(stream))); }); }

int32_t spk_alloc(void **ptr, int64_t bytes)
{
    return guard([&] {
        SPK_ARG(bytes >= 0, "Allocation size can't be negative.");
        Cache &c = cache();
        std::lock_guard<std::mutex> lock(c.mu);
        auto it = c.free.find(bytes);
        void *p = nullptr;
        if (it != c.free.end())
        {
            p = it->second;
            c.free.erase(it);
        }
        else
        {
            hipError_t e = hipMalloc(&p, size_t(bytes > 0 ? bytes : 1));
            if (e != hipSuccess)
            {
                // Out of memory: release the cache and try once more
                // (src/spyker/cuda/memory.cu:67-83 frees cached blocks under pressure).
                (void)hipGetLastError();
                SPK_HIP(hipDeviceSynchronize());
                for (auto &kv : c.free) (void)hipFree(kv.second);
                c.free.clear();
                c.cached_bytes = 0;
                SPK_HIP(hipMalloc(&p, size_t(bytes > 0 ? bytes : 1)));
            }
            c.cached_bytes += bytes;
        }
        c.taken[p] = bytes;
        c.used_bytes += bytes;
        *ptr = p;
    });
}

int32_t spk_dealloc(void *ptr)
{
    return guard([&] {
        if (!ptr) return;
        Cache &c = cache();
        std::lock_guard<std::mutex> lock(c.mu);
        auto it = c.taken.find(ptr);
        SPK_ARG(it != c.taken.end(), "Pointer was not allocated by this library.");
        c.used_bytes -= it->second;
        c.free.emplace(it->second, ptr);
        c.taken.erase(it);
    });
}

int32_t spk_copy(const void *src, void *dst, int64_t bytes, void *stream)
{
    return guard([&] {
        SPK_ARG(bytes >= 0, "Copy size can't be negative.");
        if (bytes == 0) return;
        SPK_ARG(src && dst, "Copy buffers can't be null.");
        SPK_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDefault, S(stream)));
    });
}

int32_t spk_memory_taken(int64_t *bytes)
{
    return guard([&] {
        Cache &c = cache();
        std::lock_guard<std::mutex> lock(c.mu);
        *bytes = c.cached_bytes;
    });
}

int32_t spk_memory_used(int64_t *bytes)
{
    return guard([&] {
        Cache &c = cache();
        std::lock_guard<std::mutex> lock(c.mu);
        *bytes = c.used_bytes;
    });
}

int32_t spk_cache_clear(void)
{
    return guard([&] {
        Cache &c = cache();
        std::lock_guard<std::mutex> lock(c.mu);
        SPK_HIP(hipDeviceSynchronize());
        for (auto &kv : c.free)
        {
            SPK_HIP(hipFree(kv.second));
            c.cached_bytes -= kv.first;
        }
        c.free.clear();
        clear_plan_cache();
        clear_workspaces();
    });
}
}
