// Elementwise stages: threshold, rank fire, cast, fill, pad.
//
// threshold  src/spyker/cpu/fire.cpp:41-57   (x <= th ? value : x)
// rank fire  src/spyker/cpu/fire.cpp:67-82   (x > th)
// pad        src/spyker/cpu/pool.cpp:41-51
//
// Each thread handles 16 bytes of the wider operand so every load and store
// is a dwordx4; the tail is handled element by element.
#include "common.hpp"

namespace spk
{
namespace
{
template <typename T>
__global__ void threshold_kernel(T *__restrict__ x, int64_t n, T th, T value)
{
    constexpr int V = 16 / sizeof(T);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x * V;
    for (int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * V; i < n; i += stride)
    {
        if (i + V <= n && (reinterpret_cast<uintptr_t>(x + i) & 15) == 0)
        {
            T v[V];
            *reinterpret_cast<uint4 *>(v) = *reinterpret_cast<const uint4 *>(x + i);
#pragma unroll
            for (int j = 0; j < V; ++j) v[j] = (v[j] <= th) ? value : v[j];
            *reinterpret_cast<uint4 *>(x + i) = *reinterpret_cast<uint4 *>(v);
        }
        else
            for (int64_t j = i; j < n && j < i + V; ++j) x[j] = (x[j] <= th) ? value : x[j];
    }
}

template <typename I, typename O>
__global__ void fire_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n, I th)
{
    constexpr int V = 16;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x * V;
    for (int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * V; i < n; i += stride)
    {
        if (i + V <= n && (reinterpret_cast<uintptr_t>(in + i) & 15) == 0 &&
            (reinterpret_cast<uintptr_t>(out + i) & 15) == 0)
        {
            I v[V];
            O o[V];
#pragma unroll
            for (int j = 0; j < V * int(sizeof(I)) / 16; ++j)
                reinterpret_cast<uint4 *>(v)[j] = reinterpret_cast<const uint4 *>(in + i)[j];
#pragma unroll
            for (int j = 0; j < V; ++j) o[j] = O(v[j] > th);
#pragma unroll
            for (int j = 0; j < V * int(sizeof(O)) / 16; ++j)
                reinterpret_cast<uint4 *>(out + i)[j] = reinterpret_cast<uint4 *>(o)[j];
        }
        else
            for (int64_t j = i; j < n && j < i + V; ++j) out[j] = O(in[j] > th);
    }
}

template <typename I, typename O>
__global__ void cast_kernel(const I *__restrict__ in, O *__restrict__ out, int64_t n)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = O(in[i]);
}

template <typename T>
__global__ void pad_kernel(const T *__restrict__ in, T *__restrict__ out, int64_t planes, int64_t h, int64_t w,
                           int64_t pt, int64_t pl, int64_t oh, int64_t ow, T value)
{
    const int64_t total = planes * oh * ow;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
    {
        int64_t x = i % ow, y = (i / ow) % oh, p = i / (ow * oh);
        int64_t iy = y - pt, ix = x - pl;
        out[i] = (iy >= 0 && iy < h && ix >= 0 && ix < w) ? in[(p * h + iy) * w + ix] : value;
    }
}

template <typename T>
void threshold_t(void *x, int64_t n, double th, double value, hipStream_t s)
{
    threshold_kernel<T><<<flat_grid(n, 256 * (16 / sizeof(T))), 256, 0, s>>>((T *)x, n, T(th), T(value));
    SPK_LAUNCHED();
}

template <typename I, typename O>
void fire_t(const void *in, void *out, int64_t n, double th, hipStream_t s)
{
    fire_kernel<I, O><<<flat_grid(n, 256 * 16), 256, 0, s>>>((const I *)in, (O *)out, n, I(th));
    SPK_LAUNCHED();
}

template <typename I, typename O>
void cast_t(const void *in, void *out, int64_t n, hipStream_t s)
{
    cast_kernel<I, O><<<flat_grid(n, 256), 256, 0, s>>>((const I *)in, (O *)out, n);
    SPK_LAUNCHED();
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ x, int64_t n, T value)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = value;
}

// Calls f with a value of the element type `type` names (the reference's
// IfType switch, src/spyker/base.h:583-776).
template <typename F>
void with_type(int32_t type, F &&f)
{
    switch (type)
    {
    case SPK_I8: return f(int8_t{});
    case SPK_I16: return f(int16_t{});
    case SPK_I32: return f(int32_t{});
    case SPK_I64: return f(int64_t{});
    case SPK_U8: return f(uint8_t{});
    case SPK_U16: return f(uint16_t{});
    case SPK_U32: return f(uint32_t{});
    case SPK_U64: return f(uint64_t{});
    case SPK_F16: return f(_Float16{});
    case SPK_F32: return f(float{});
    case SPK_F64: return f(double{});
    default: SPK_ARG(false, "Data type " << type << " is not recognized.");
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_threshold(void *x, int32_t type, int64_t n, double threshold, double value, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0, "Input size can't be negative.");
        if (n == 0) return;
        switch (type)
        {
        case SPK_F32: return threshold_t<float>(x, n, threshold, value, S(stream));
        case SPK_F16: return threshold_t<_Float16>(x, n, threshold, value, S(stream));
        case SPK_F64: return threshold_t<double>(x, n, threshold, value, S(stream));
        default: SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "Threshold: data type " << type_name(type) << " is not supported.");
        }
    });
}

int32_t spk_rank_fire(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, double threshold,
                      void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0, "Input size can't be negative.");
        if (n == 0) return;
        hipStream_t s = S(stream);
        if (in_type == SPK_F32 && out_type == SPK_U8) return fire_t<float, uint8_t>(in, out, n, threshold, s);
        if (in_type == SPK_F32 && out_type == SPK_F32) return fire_t<float, float>(in, out, n, threshold, s);
        if (in_type == SPK_F16 && out_type == SPK_U8) return fire_t<_Float16, uint8_t>(in, out, n, threshold, s);
        if (in_type == SPK_F64 && out_type == SPK_U8) return fire_t<double, uint8_t>(in, out, n, threshold, s);
        SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED,
                    "Fire: data types " << type_name(in_type) << " -> " << type_name(out_type) << " are not supported.");
    });
}

int32_t spk_cast(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0, "Input size can't be negative.");
        if (n == 0) return;
        hipStream_t s = S(stream);
        with_type(in_type, [&](auto a) {
            with_type(out_type, [&](auto b) { cast_t<decltype(a), decltype(b)>(in, out, n, s); });
        });
    });
}

int32_t spk_fill(void *x, int32_t type, int64_t n, double value, void *stream)
{
    return guard([&] {
        SPK_ARG(n >= 0, "Input size can't be negative.");
        if (n == 0) return;
        hipStream_t s = S(stream);
        with_type(type, [&](auto a) {
            using T = decltype(a);
            fill_kernel<T><<<flat_grid(n, 256), 256, 0, s>>>((T *)x, n, T(value));
            SPK_LAUNCHED();
        });
    });
}

int32_t spk_pad(const void *in, void *out, int32_t type, int64_t planes, int64_t h, int64_t w, int64_t pt, int64_t pl,
                int64_t pb, int64_t pr, double value, void *stream)
{
    return guard([&] {
        SPK_ARG(planes >= 0 && h > 0 && w > 0, "Pad: input shape is invalid.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        int64_t oh = h + pt + pb, ow = w + pl + pr, total = planes * oh * ow;
        if (total == 0) return;
        hipStream_t s = S(stream);
        if (type == SPK_U8)
            pad_kernel<uint8_t><<<flat_grid(total, 256), 256, 0, s>>>((const uint8_t *)in, (uint8_t *)out, planes, h,
                                                                       w, pt, pl, oh, ow, uint8_t(value));
        else if (type == SPK_F32)
            pad_kernel<float><<<flat_grid(total, 256), 256, 0, s>>>((const float *)in, (float *)out, planes, h, w, pt,
                                                                     pl, oh, ow, float(value));
        else if (type == SPK_F16)
            pad_kernel<_Float16><<<flat_grid(total, 256), 256, 0, s>>>((const _Float16 *)in, (_Float16 *)out, planes,
                                                                        h, w, pt, pl, oh, ow, _Float16(value));
        else if (type == SPK_F64)
            pad_kernel<double><<<flat_grid(total, 256), 256, 0, s>>>((const double *)in, (double *)out, planes, h, w,
                                                                      pt, pl, oh, ow, value);
        else
            SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "Pad: data type " << type_name(type) << " is not supported.");
        SPK_LAUNCHED();
    });
}
}
