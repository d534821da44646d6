// Shared host/device helpers for the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <string>

#include "spyker_hip.h"

namespace spk
{
// ---------------------------------------------------------------- errors --

struct Error : std::runtime_error
{
    int32_t code;
    Error(int32_t code, const std::string &msg) : std::runtime_error(msg), code(code) {}
};

void set_last_error(const std::string &msg);

#define SPK_REQUIRE(cond, code, msg)                                     \
    do                                                                   \
    {                                                                    \
        if (!(cond))                                                     \
        {                                                                \
            std::ostringstream _s;                                       \
            _s << msg;                                                   \
            throw ::spk::Error(code, _s.str());                          \
        }                                                                \
    } while (0)

#define SPK_ARG(cond, msg) SPK_REQUIRE(cond, SPK_ERR_ARGUMENT, msg)
// outside what a device path implements (a caller may route the call elsewhere)
#define SPK_UNSUP(cond, msg) SPK_REQUIRE(cond, SPK_ERR_UNSUPPORTED, msg)

#define SPK_HIP(expr)                                                                                       \
    do                                                                                                      \
    {                                                                                                       \
        hipError_t _e = (expr);                                                                             \
        SPK_REQUIRE(_e == hipSuccess, SPK_ERR_RUNTIME, #expr << ": " << hipGetErrorString(_e));             \
    } while (0)

#define SPK_LAUNCHED() SPK_HIP(hipGetLastError())

// Runs a C-ABI body, mapping exceptions to a status code and the message.
template <typename F>
int32_t guard(F &&f)
{
    try
    {
        f();
        return SPK_OK;
    }
    catch (const Error &e)
    {
        set_last_error(e.what());
        return e.code;
    }
    catch (const std::exception &e)
    {
        set_last_error(e.what());
        return SPK_ERR_RUNTIME;
    }
}

inline hipStream_t S(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

const char *type_name(int32_t type);

// Frees the layer plans the uniform fused path keeps per shape (region.hip).
void clear_plan_cache();
// Compute units of the current device (queried once per device).
int cu_count();
// A grow-only device workspace of at least `bytes` for the calls of one
// entry point (`owner`) on one stream of the current device.  Calls on a
// stream run in order, so they share it without waiting; it grows only
// outside a stream capture (a capture that needs more is refused, so warm
// the call up first), and spk_cache_clear frees it.
void *stream_workspace(int owner, hipStream_t stream, int64_t bytes);
void clear_workspaces();
// One draw of the shared host generator (Generator() of the reference).
uint64_t generator_draw();
// [rows][bins] Poisson interval table (Core::poisson_create), host memory.
void poisson_rows(uint16_t *table, int64_t rows, int64_t bins);

// --------------------------------------------------------------- launch --

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid size for a flat loop of n items, b per block, capped so grid-stride
// loops take over beyond ~16 waves of blocks per CU.
inline dim3 flat_grid(int64_t n, int64_t per_block)
{
    int64_t g = cdiv(n, per_block);
    if (g > 256 * 64) g = 256 * 64;
    if (g < 1) g = 1;
    return dim3(unsigned(g));
}

// Per-call workspace from the stream-ordered pool (hipMallocAsync on the
// call's stream, hipFreeAsync in the destructor), so a block is reused only
// after the work enqueued on that stream before the release has finished.
struct Scratch
{
    void *ptr = nullptr;
    int64_t bytes = 0;
    hipStream_t stream = nullptr;
    Scratch(int64_t bytes, hipStream_t stream);
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    template <typename T>
    T *as() const
    {
        return reinterpret_cast<T *>(ptr);
    }
};

// ---------------------------------------------------------------- device --

// Lexicographic (value descending, index ascending) maximum: the first
// maximum of an ascending scan with strict '>' (std::max_element and the
// reference's inhibit loop both pick the lowest index among equal maxima).
struct ArgMax
{
    float v;
    int32_t i;
};

__device__ __forceinline__ bool better(float va, int32_t ia, float vb, int32_t ib)
{
    return va > vb || (va == vb && ia < ib);
}

__device__ __forceinline__ ArgMax wave_argmax(ArgMax a)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        float v = __shfl_xor(a.v, off, 64);
        int32_t i = __shfl_xor(a.i, off, 64);
        if (better(v, i, a.v, a.i)) a.v = v, a.i = i;
    }
    return a;
}

// Block-wide argmax over blockDim.x threads (a multiple of 64, <= 1024);
// every thread receives the result.
__device__ __forceinline__ ArgMax block_argmax(ArgMax a, ArgMax *lds /* [16] */)
{
    a = wave_argmax(a);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = a;
    __syncthreads();
    ArgMax r = lds[0];
    for (int i = 1; i < waves; ++i)
        if (better(lds[i].v, lds[i].i, r.v, r.i)) r = lds[i];
    return r;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        float o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        float o = __shfl_xor(v, off, 64);
        v = v < o ? o : v;
    }
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int32_t wave_sum_i32(int32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide min and max in the reference's std::min/std::max form
// (min = v < min ? v : min; max = max < v ? v : max).
__device__ __forceinline__ void block_minmax(const float *__restrict__ x, int64_t len, float &mn, float &mx, float *lds /*[32]*/)
{
    float a = INFINITY, b = -INFINITY;  // neutral for the reductions below
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x)
    {
        float v = x[i];
        a = v < a ? v : a, b = b < v ? v : b;
    }
    a = wave_min(a), b = wave_max(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = a, lds[16 + wave] = b;
    __syncthreads();
    mn = lds[0], mx = lds[16];
    for (int i = 1; i < waves; ++i) mn = lds[i] < mn ? lds[i] : mn, mx = mx < lds[16 + i] ? lds[16 + i] : mx;
}

}  // namespace spk
