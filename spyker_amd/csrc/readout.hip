// Readout after the layer: first-spike times and back (src/spyker/cpu/time.cpp,
// src/spyker/cuda/time.cu).  Elementwise over (sample, element); a thread
// owns one element and walks the steps, so every step's loads or stores are
// coalesced across the threads of a wave.
#include <limits>
#include <type_traits>

#include "common.hpp"

namespace spk
{
namespace
{
// The reference converts the threshold Scalar to the input type; integer
// types take it truncated toward zero, clamped into their range.
template <typename I>
__host__ __device__ I to_input(double v)
{
    if constexpr (std::is_integral<I>::value)
    {
        const double lo = double(std::numeric_limits<I>::min()), hi = double(std::numeric_limits<I>::max());
        return I(v < lo ? lo : v > hi ? hi : v);
    }
    else
        return I(v);
}

// rank gather (time.cpp:40-64): the first step with in > threshold, T where none.
template <typename I, typename O>
__global__ void rank_gather_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n, int32_t t, int64_t len,
                                   I th)
{
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * len;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / len, j = idx % len;
        const I *x = in + i * t * len + j;
        int32_t f = t;
        for (int32_t s = 0; s < t; ++s)
            if (x[int64_t(s) * len] > th)
            {
                f = s;
                break;
            }
        out[idx] = O(f);
    }
}

// rank scatter (time.cpp:66-84): step s spikes where the unsigned step index
// s >= the element's time, compared as the reference's C++ does.
template <typename I>
__device__ __forceinline__ bool reached(uint32_t s, I v)
{
    if constexpr (std::is_floating_point<I>::value)
        return float(s) >= v;
    else
        return s >= uint32_t(v);  // negative times compare as large unsigned values
}

template <typename I, typename O>
__global__ void rank_scatter_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n, int64_t len, int32_t t)
{
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * len;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / len, j = idx % len;
        const I v = in[idx];
        O *o = out + i * t * len + j;
        for (int32_t s = 0; s < t; ++s) o[int64_t(s) * len] = O(reached(uint32_t(s), v) ? 1 : 0);
    }
}

// rate gather (time.cpp:86-92): the last step, converted.
template <typename I, typename O>
__global__ void rate_gather_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n, int32_t t, int64_t len)
{
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * len;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / len, j = idx % len;
        const I v = in[(i * t + t - 1) * len + j];
        if constexpr (std::is_integral<O>::value && std::is_floating_point<I>::value)
        {
            const float lo = float(std::numeric_limits<O>::min()), hi = float(std::numeric_limits<O>::max());
            out[idx] = O(v < lo ? lo : v > hi ? hi : v);
        }
        else
            out[idx] = O(v);
    }
}

// rate fire (fire.cpp:84-114; cuda/fire.cu:82-96): per element over the steps,
// value + threshold < x counts a spike (in the output type) and keeps x.  A
// thread walks W consecutive elements through the steps (W = 4 when the row
// length allows it, so each step is one 4-wide load and store per thread).
template <typename I, typename O, int W>
__global__ void rate_fire_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n, int32_t t, int64_t len,
                                 I th)
{
    const int64_t groups = len / W;
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * groups;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / groups, j = (idx % groups) * W;
        const I *x = in + i * t * len + j;
        O *o = out + i * t * len + j;
        O spikes[W];
        I value[W];
#pragma unroll
        for (int k = 0; k < W; ++k) spikes[k] = O(0), value[k] = I(0);
        for (int32_t s = 0; s < t; ++s)
        {
            I v[W];
            O r[W];
            if constexpr (W == 4 && sizeof(I) == 4)
                *reinterpret_cast<uint4 *>(v) = *reinterpret_cast<const uint4 *>(x + int64_t(s) * len);
            else
#pragma unroll
                for (int k = 0; k < W; ++k) v[k] = x[int64_t(s) * len + k];
#pragma unroll
            for (int k = 0; k < W; ++k)
            {
                if (value[k] + th < v[k]) spikes[k] = O(spikes[k] + O(1)), value[k] = v[k];
                r[k] = spikes[k];
            }
            if constexpr (W == 4 && sizeof(O) == 1)
                *reinterpret_cast<uint32_t *>(o + int64_t(s) * len) = *reinterpret_cast<const uint32_t *>(r);
            else if constexpr (W == 4 && sizeof(O) == 4)
                *reinterpret_cast<uint4 *>(o + int64_t(s) * len) = *reinterpret_cast<const uint4 *>(r);
            else
#pragma unroll
                for (int k = 0; k < W; ++k) o[int64_t(s) * len + k] = r[k];
        }
    }
}

// Calls f with a value of the C type of `type` (U8, U16, I32 or F32).
template <typename F>
void with_type(int32_t type, const char *what, F &&f)
{
    switch (type)
    {
    case SPK_U8: return f(uint8_t());
    case SPK_U16: return f(uint16_t());
    case SPK_I32: return f(int32_t());
    case SPK_F32: return f(float());
    default: SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, what << ": data type " << type_name(type) << " is not supported.");
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_rank_gather(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                        int64_t len, double threshold, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0 && t >= 1 && len >= 0, "Gather: input shape is invalid.");
        SPK_ARG(t < INT32_MAX, "Gather: too many time steps.");
        hipStream_t s = S(stream);
        with_type(in_type, "Gather", [&](auto i0) {
            using I = decltype(i0);
            with_type(out_type, "Gather", [&](auto o0) {
                using O = decltype(o0);
                if (n * len == 0) return;
                rank_gather_kernel<I, O><<<flat_grid(n * len, 256), 256, 0, s>>>(
                    (const I *)in, (O *)out, n, int32_t(t), len, to_input<I>(threshold));
                SPK_LAUNCHED();
            });
        });
    });
}

int32_t spk_rank_scatter(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                         int64_t t, void *stream)
{
    return guard([&] {
        SPK_ARG(t > 0, "Input time can't be less than 1.");
        SPK_ARG(n >= 0 && len >= 0 && t < INT32_MAX, "Scatter: input shape is invalid.");
        hipStream_t s = S(stream);
        with_type(in_type, "Scatter", [&](auto i0) {
            using I = decltype(i0);
            with_type(out_type, "Scatter", [&](auto o0) {
                using O = decltype(o0);
                if (n * len == 0) return;
                rank_scatter_kernel<I, O><<<flat_grid(n * len, 256), 256, 0, s>>>((const I *)in, (O *)out, n, len,
                                                                                  int32_t(t));
                SPK_LAUNCHED();
            });
        });
    });
}

int32_t spk_rate_fire(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                      int64_t len, double threshold, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0 && t >= 1 && len >= 0 && t < INT32_MAX, "Fire: input shape is invalid.");
        hipStream_t s = S(stream);
        with_type(in_type, "Fire", [&](auto i0) {
            using I = decltype(i0);
            with_type(out_type, "Fire", [&](auto o0) {
                using O = decltype(o0);
                if (n * len == 0) return;
                // 4-wide when rows keep 16-byte input and 4-element output alignment
                const bool wide = len % 4 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                                  reinterpret_cast<uintptr_t>(out) % (4 * sizeof(O)) == 0;
                if (wide)
                    rate_fire_kernel<I, O, 4><<<flat_grid(n * len / 4, 256), 256, 0, s>>>(
                        (const I *)in, (O *)out, n, int32_t(t), len, to_input<I>(threshold));
                else
                    rate_fire_kernel<I, O, 1><<<flat_grid(n * len, 256), 256, 0, s>>>(
                        (const I *)in, (O *)out, n, int32_t(t), len, to_input<I>(threshold));
                SPK_LAUNCHED();
            });
        });
    });
}

int32_t spk_rate_gather(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                        int64_t len, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0 && t >= 1 && len >= 0 && t < INT32_MAX, "Gather: input shape is invalid.");
        hipStream_t s = S(stream);
        with_type(in_type, "Gather", [&](auto i0) {
            using I = decltype(i0);
            with_type(out_type, "Gather", [&](auto o0) {
                using O = decltype(o0);
                if (n * len == 0) return;
                rate_gather_kernel<I, O><<<flat_grid(n * len, 256), 256, 0, s>>>((const I *)in, (O *)out, n,
                                                                                 int32_t(t), len);
                SPK_LAUNCHED();
            });
        });
    });
}
}
