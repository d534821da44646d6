// Pooling, lateral inhibition and winner-take-all for the plain path
// (F16, F32 and F64 values; U8 spikes where the reference takes them).
//
// pool     src/spyker/cpu/pool.cpp:53-86 (zero pad, window max seeded with
//          the lowest value, scanned row-major so signed zeros resolve as
//          in the reference)
// inhibit  src/spyker/cpu/inhibit.cpp:41-121
// convwta  src/spyker/cpu/wta.cpp:44-65,128-182
// fcwta    src/spyker/cpu/wta.cpp:87-126
#include "common.hpp"

namespace spk
{
namespace
{
template <typename T>
struct Lowest;
template <>
struct Lowest<uint8_t>
{
    static __device__ uint8_t v() { return 0; }
};
template <>
struct Lowest<float>
{
    static __device__ float v() { return -FLT_MAX; }
};
template <>
struct Lowest<double>
{
    static __device__ double v() { return -DBL_MAX; }
};
// Limits<half>::min() (base.h:772): the smallest positive normal, so a window
// of zeros or negatives pools to 2^-14, as the reference's does
template <>
struct Lowest<_Float16>
{
    static __device__ _Float16 v() { return _Float16(6.103515625e-05f); }
};

template <typename T>
__global__ void pool_kernel(const T *__restrict__ in, T *__restrict__ out, int64_t planes, int32_t h, int32_t w,
                            int32_t oh, int32_t ow, int32_t kh, int32_t kw, int32_t sh, int32_t sw, int32_t pt,
                            int32_t pl)
{
    const int64_t total = planes * oh * ow;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
    {
        const int32_t x = int32_t(i % ow), y = int32_t((i / ow) % oh);
        const T *src = in + (i / (int64_t(ow) * oh)) * int64_t(h) * w;
        T m = Lowest<T>::v();
        for (int32_t a = 0; a < kh; ++a)
        {
            const int32_t iy = y * sh + a - pt;
            for (int32_t b = 0; b < kw; ++b)
            {
                const int32_t ix = x * sw + b - pl;
                T v = (iy >= 0 && iy < h && ix >= 0 && ix < w) ? src[iy * w + ix] : T(0);
                m = m < v ? v : m;
            }
        }
        out[i] = m;
    }
}

// One thread per (sample, pixel); consecutive threads take consecutive
// pixels, so each (t, c) row is read coalesced.  Values compare and scale in
// their own type T, the threshold converted to T (inhibit.cpp:41-121).
template <typename T>
__global__ void inhibit_kernel(T *__restrict__ x, int64_t n, int32_t t, int32_t c, int64_t hw, T th)
{
    const int64_t total = n * hw;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
    {
        const int64_t p = i % hw;
        T *xs = x + (i / hw) * int64_t(t) * c * hw + p;
        // cnt = #{steps whose channel maximum exceeds th}
        int32_t cnt = 0;
        for (int32_t s = 0; s < t; ++s)
        {
            const T *row = xs + int64_t(s) * c * hw;
            T m = row[0];
            for (int32_t j = 1; j < c; ++j)
            {
                T v = row[int64_t(j) * hw];
                m = v > m ? v : m;
            }
            cnt += m > th;
        }
        // kept channel: first maximum at step T - max(cnt, 1)
        const int32_t ts = t - (cnt > 1 ? cnt : 1);
        const T *row = xs + int64_t(ts) * c * hw;
        T m = row[0];
        int32_t keep = 0;
        for (int32_t j = 1; j < c; ++j)
        {
            T v = row[int64_t(j) * hw];
            if (v > m) m = v, keep = j;
        }
        for (int32_t s = 0; s < t; ++s)
            for (int32_t j = 0; j < c; ++j)
            {
                T &v = xs[(int64_t(s) * c + j) * hw];
                v = v * (j == keep ? T(1) : T(0));
            }
    }
}

// The reference's Limits<T>::min() (base.h:761-774): the lowest F32 or F64,
// but for F16 numeric_limits<half>::min(), the smallest positive normal
// 2^-14.  Unfired neurons hold it and the selection stops at it.
template <typename T>
struct WtaLowest;
template <>
struct WtaLowest<float>
{
    static __device__ float v() { return -FLT_MAX; }
};
template <>
struct WtaLowest<double>
{
    static __device__ double v() { return -DBL_MAX; }
};
template <>
struct WtaLowest<_Float16>
{
    static __device__ _Float16 v() { return _Float16(6.103515625e-05f); }
};

// Values are computed in T and held widened (exactly) in R: F32 for F16 and
// F32 values, F64 for F64.
template <typename T>
using WtaR = typename std::conditional<std::is_same<T, double>::value, double, float>::type;

// a / b correctly rounded in T (an F16 quotient through the correctly
// rounded F32 one: 24 bits hold the F16 result twice over, so the second
// rounding cannot change it)
template <typename T>
__device__ __forceinline__ T wta_div(T a, T b)
{
    if constexpr (std::is_same<T, _Float16>::value)
        return _Float16(float(a) / float(b));
    else
        return a / b;
}

// |v - lowest| < Epsilon (1e-5f) with the difference taken in T (wta.cpp:93,136)
template <typename T>
__device__ __forceinline__ bool wta_stop(WtaR<T> v)
{
    const T d = T(v) - WtaLowest<T>::v();
    return (d < T(0) ? -d : d) < 1e-5f;
}

template <typename R>
struct ArgMaxR
{
    R v;
    int32_t i;
};

template <typename R>
__device__ __forceinline__ ArgMaxR<R> block_argmax_r(ArgMaxR<R> a, ArgMaxR<R> *lds /* [16] */)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        R v = __shfl_xor(a.v, off, 64);
        int32_t i = __shfl_xor(a.i, off, 64);
        if (v > a.v || (v == a.v && i < a.i)) a.v = v, a.i = i;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = a;
    __syncthreads();
    ArgMaxR<R> r = lds[0];
    for (int i = 1; i < waves; ++i)
        if (lds[i].v > r.v || (lds[i].v == r.v && lds[i].i < r.i)) r = lds[i];
    return r;
}

// Block-wide min and max of T values in the reference's std::min/std::max
// form, widened to R (exact).
template <typename T>
__device__ __forceinline__ void block_minmax_t(const T *__restrict__ x, int64_t len, WtaR<T> &mn, WtaR<T> &mx,
                                               WtaR<T> *lds /*[32]*/)
{
    using R = WtaR<T>;
    R a = R(INFINITY), b = R(-INFINITY);
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x)
    {
        const R v = R(x[i]);
        a = v < a ? v : a, b = b < v ? v : b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const R o = __shfl_xor(a, off, 64), q = __shfl_xor(b, off, 64);
        a = o < a ? o : a, b = b < q ? q : b;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = a, lds[16 + wave] = b;
    __syncthreads();
    mn = lds[0], mx = lds[16];
    for (int i = 1; i < waves; ++i) mn = lds[i] < mn ? lds[i] : mn, mx = mx < lds[16 + i] ? lds[16 + i] : mx;
}

// Value of each neuron: (T-1-t0) + (x[t0]-min)/diff in T at the earliest step
// t0 above threshold, else the lowest value (wta.cpp:44-64).
template <typename T>
__device__ void wta_values(const T *__restrict__ x, int32_t t, int64_t len, T th, T mn, T diff,
                           WtaR<T> *__restrict__ temp)
{
    for (int64_t j = threadIdx.x; j < len; j += blockDim.x)
    {
        T v = WtaLowest<T>::v();
        for (int32_t s = 0; s < t; ++s)
        {
            const T xv = x[int64_t(s) * len + j];
            if (xv > th)
            {
                v = T(t - s - 1) + wta_div<T>(T(xv - mn), diff);
                break;
            }
        }
        temp[j] = WtaR<T>(v);
    }
}

// Integer input: T - first spiking step, 0 if none (wta.cpp:66-85).
__device__ void wta_values_u8(const uint8_t *__restrict__ x, int32_t t, int64_t len, float *__restrict__ temp)
{
    for (int64_t j = threadIdx.x; j < len; j += blockDim.x)
    {
        float v = 0.0f;
        for (int32_t s = 0; s < t; ++s)
            if (x[int64_t(s) * len + j])
            {
                v = float(t - s);
                break;
            }
        temp[j] = v;
    }
}

// k rounds of: first maximum, stop test, emit, mask the box in every channel
// and the winner's channel (wta.cpp:128-150).  T = uint8_t: integer values,
// stop at 0.
template <typename T>
__device__ void conv_select(WtaR<T> *__restrict__ temp, int32_t c, int32_t h, int32_t w, int32_t ry, int32_t rx,
                            int32_t count, spk_winner *__restrict__ win, int64_t *__restrict__ cnt,
                            ArgMaxR<WtaR<T>> *lds)
{
    using R = WtaR<T>;
    constexpr bool Integer = std::is_same<T, uint8_t>::value;
    const int64_t hw = int64_t(h) * w, len = hw * c;
    R lowest;
    if constexpr (Integer)
        lowest = R(0);
    else
        lowest = R(WtaLowest<T>::v());
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMaxR<R> a{R(-INFINITY), INT32_MAX};
        for (int64_t j = threadIdx.x; j < len; j += blockDim.x)
        {
            const R v = temp[j];
            if (v > a.v || (v == a.v && int32_t(j) < a.i)) a.v = v, a.i = int32_t(j);
        }
        a = block_argmax_r(a, lds);
        bool stop;
        if constexpr (Integer)
            stop = a.v == R(0);
        else
            stop = wta_stop<T>(a.v);
        if (stop) break;
        const int32_t z = int32_t(a.i / hw), y = int32_t((a.i % hw) / w), x = int32_t((a.i % hw) % w);
        if (threadIdx.x == 0) win[k] = spk_winner{0, 0, z, y, x};
        const int32_t sj = max(0, y - ry), ej = min(h, y + ry + 1);
        const int32_t sk = max(0, x - rx), ek = min(w, x + rx + 1);
        const int32_t bh = ej - sj, bw = ek - sk;
        for (int64_t j = threadIdx.x; j < int64_t(c) * bh * bw; j += blockDim.x)
        {
            int32_t ch = int32_t(j / (bh * bw)), r = int32_t((j / bw) % bh), q = int32_t(j % bw);
            temp[int64_t(ch) * hw + int64_t(sj + r) * w + (sk + q)] = lowest;
        }
        for (int64_t j = threadIdx.x; j < hw; j += blockDim.x) temp[int64_t(z) * hw + j] = lowest;
        __syncthreads();
    }
    if (threadIdx.x == 0) *cnt = k;
}

template <typename T>
__global__ void __launch_bounds__(1024) convwta_kernel(const T *__restrict__ x, int32_t t, int32_t c, int32_t h,
                                                       int32_t w, int32_t ry, int32_t rx, int32_t count, T th,
                                                       WtaR<T> *__restrict__ temp, spk_winner *__restrict__ win,
                                                       int64_t *__restrict__ cnt)
{
    using R = WtaR<T>;
    __shared__ R mm[32];
    __shared__ ArgMaxR<R> am[16];
    const int64_t s = blockIdx.x, len = int64_t(c) * h * w;
    const T *xs = x + s * t * len;
    R *ts = temp + s * len;
    R mn, mx;
    block_minmax_t(xs, int64_t(t) * len, mn, mx, mm);
    const T tmn = T(mn), diff = T(T(T(mx) - tmn) * T(2));
    wta_values(xs, t, len, th, tmn, diff, ts);
    __syncthreads();
    conv_select<T>(ts, c, h, w, ry, rx, count, win + s * count, cnt + s, am);
}

__global__ void __launch_bounds__(1024) convwta_u8_kernel(const uint8_t *__restrict__ x, int32_t t, int32_t c,
                                                          int32_t h, int32_t w, int32_t ry, int32_t rx,
                                                          int32_t count, float *__restrict__ temp,
                                                          spk_winner *__restrict__ win, int64_t *__restrict__ cnt)
{
    __shared__ ArgMaxR<float> am[16];
    const int64_t s = blockIdx.x, len = int64_t(c) * h * w;
    float *ts = temp + s * len;
    wta_values_u8(x + s * t * len, t, len, ts);
    __syncthreads();
    conv_select<uint8_t>(ts, c, h, w, ry, rx, count, win + s * count, cnt + s, am);
}

template <typename T>
__global__ void __launch_bounds__(1024) fcwta_kernel(const T *__restrict__ x, int32_t t, int32_t f, int32_t radius,
                                                     int32_t count, T th, WtaR<T> *__restrict__ temp,
                                                     spk_winner *__restrict__ win, int64_t *__restrict__ cnt)
{
    using R = WtaR<T>;
    __shared__ R mm[32];
    __shared__ ArgMaxR<R> am[16];
    const int64_t s = blockIdx.x;
    const T *xs = x + s * t * int64_t(f);
    R *ts = temp + s * int64_t(f);
    R mn, mx;
    block_minmax_t(xs, int64_t(t) * f, mn, mx, mm);
    const T tmn = T(mn), diff = T(T(T(mx) - tmn) * T(2));
    wta_values(xs, t, f, th, tmn, diff, ts);
    __syncthreads();
    const R lowest = R(WtaLowest<T>::v());
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMaxR<R> a{R(-INFINITY), INT32_MAX};
        for (int32_t j = threadIdx.x; j < f; j += blockDim.x)
            if (ts[j] > a.v || (ts[j] == a.v && j < a.i)) a.v = ts[j], a.i = j;
        a = block_argmax_r(a, am);
        if (wta_stop<T>(a.v)) break;
        if (threadIdx.x == 0) win[s * count + k] = spk_winner{0, 0, 0, 0, a.i};
        const int32_t lo = max(0, a.i - radius), hi = min(f, a.i + radius + 1);
        for (int32_t j = lo + int32_t(threadIdx.x); j < hi; j += blockDim.x) ts[j] = lowest;
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[s] = k;
}

// Integer input (wta.cpp:87-126 with the U8 values): first maximum of
// T - first spiking step, stop at 0, the window is set to 0.
__global__ void __launch_bounds__(1024) fcwta_u8_kernel(const uint8_t *__restrict__ x, int32_t t, int32_t f,
                                                        int32_t radius, int32_t count, float *__restrict__ temp,
                                                        spk_winner *__restrict__ win, int64_t *__restrict__ cnt)
{
    __shared__ ArgMax am[16];
    const int64_t s = blockIdx.x;
    float *ts = temp + s * int64_t(f);
    wta_values_u8(x + s * t * int64_t(f), t, f, ts);
    __syncthreads();
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMax a{-INFINITY, INT32_MAX};
        for (int32_t j = threadIdx.x; j < f; j += blockDim.x)
            if (better(ts[j], j, a.v, a.i)) a.v = ts[j], a.i = j;
        a = block_argmax(a, am);
        if (a.v == 0.0f) break;
        if (threadIdx.x == 0) win[s * count + k] = spk_winner{0, 0, 0, 0, a.i};
        const int32_t lo = max(0, a.i - radius), hi = min(f, a.i + radius + 1);
        for (int32_t j = lo + int32_t(threadIdx.x); j < hi; j += blockDim.x) ts[j] = 0.0f;
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[s] = k;
}

int wta_threads(int64_t len) { return len >= 64 * 1024 ? 1024 : len >= 8192 ? 512 : 256; }
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_rank_pool(const void *in, void *out, int32_t type, int64_t planes, int64_t h, int64_t w, int64_t kh,
                      int64_t kw, int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        SPK_ARG(kh >= 1 && kw >= 1, "Kernel items are incorrect.");
        SPK_ARG(sh >= 1 && sw >= 1, "Stride items are incorrect.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        const int64_t oh = (h + pt + pb - kh) / sh + 1, ow = (w + pl + pr - kw) / sw + 1;
        SPK_ARG(oh > 0 && ow > 0, "Input and kernel shapes are not compatible.");
        const int64_t total = planes * oh * ow;
        if (total == 0) return;
        hipStream_t s = S(stream);
        if (type == SPK_U8)
            pool_kernel<uint8_t><<<flat_grid(total, 256), 256, 0, s>>>(
                (const uint8_t *)in, (uint8_t *)out, planes, int32_t(h), int32_t(w), int32_t(oh), int32_t(ow),
                int32_t(kh), int32_t(kw), int32_t(sh), int32_t(sw), int32_t(pt), int32_t(pl));
        else if (type == SPK_F32)
            pool_kernel<float><<<flat_grid(total, 256), 256, 0, s>>>(
                (const float *)in, (float *)out, planes, int32_t(h), int32_t(w), int32_t(oh), int32_t(ow),
                int32_t(kh), int32_t(kw), int32_t(sh), int32_t(sw), int32_t(pt), int32_t(pl));
        else if (type == SPK_F16)
            pool_kernel<_Float16><<<flat_grid(total, 256), 256, 0, s>>>(
                (const _Float16 *)in, (_Float16 *)out, planes, int32_t(h), int32_t(w), int32_t(oh), int32_t(ow),
                int32_t(kh), int32_t(kw), int32_t(sh), int32_t(sw), int32_t(pt), int32_t(pl));
        else if (type == SPK_F64)
            pool_kernel<double><<<flat_grid(total, 256), 256, 0, s>>>(
                (const double *)in, (double *)out, planes, int32_t(h), int32_t(w), int32_t(oh), int32_t(ow),
                int32_t(kh), int32_t(kw), int32_t(sh), int32_t(sw), int32_t(pt), int32_t(pl));
        else
            SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "Pool: data type " << type_name(type) << " is not supported.");
        SPK_LAUNCHED();
    });
}

int32_t spk_rank_inhibit(void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t hw, double threshold,
                         void *stream)
{
    return guard([&] {
        SPK_REQUIRE(type == SPK_F16 || type == SPK_F32 || type == SPK_F64, SPK_ERR_UNSUPPORTED,
                    "Inhibit: data type " << type_name(type) << " is not supported.");
        SPK_ARG(n >= 0 && t >= 1 && c >= 1 && hw >= 0, "Inhibit: input shape is invalid.");
        SPK_ARG(c <= 65536 && t <= 65535, "Inhibit: channel or time count exceeds the U16 index range.");
        if (n * hw == 0) return;
        const dim3 grid = flat_grid(n * hw, 256);
        if (type == SPK_F32)
            inhibit_kernel<float><<<grid, 256, 0, S(stream)>>>((float *)x, n, int32_t(t), int32_t(c), hw,
                                                                float(threshold));
        else if (type == SPK_F16)
            inhibit_kernel<_Float16><<<grid, 256, 0, S(stream)>>>((_Float16 *)x, n, int32_t(t), int32_t(c), hw,
                                                                   _Float16(threshold));
        else
            inhibit_kernel<double><<<grid, 256, 0, S(stream)>>>((double *)x, n, int32_t(t), int32_t(c), hw, threshold);
        SPK_LAUNCHED();
    });
}

int32_t spk_rank_convwta(const void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                         int64_t ry, int64_t rx, int64_t count, double threshold, spk_winner *winners,
                         int64_t *counts, void *stream)
{
    return guard([&] {
        SPK_ARG(ry >= 0 && rx >= 0, "Radius items are incorrect.");
        SPK_ARG(count > 0, "Winners count can't be less than 1.");
        SPK_ARG(n >= 0 && t >= 1 && c >= 1 && h >= 1 && w >= 1, "ConvWTA: input shape is invalid.");
        SPK_ARG(c * h * w < INT32_MAX, "ConvWTA: a sample is too large.");
        if (n == 0) return;
        const int64_t len = c * h * w;
        hipStream_t s = S(stream);
        Scratch temp(n * len * int64_t(type == SPK_F64 ? sizeof(double) : sizeof(float)), s);
        const int threads = wta_threads(int64_t(t) * len);
        const dim3 grid = dim3(unsigned(n));
        if (type == SPK_F32)
            convwta_kernel<float><<<grid, threads, 0, s>>>(
                (const float *)x, int32_t(t), int32_t(c), int32_t(h), int32_t(w), int32_t(ry), int32_t(rx),
                int32_t(count), float(threshold), temp.as<float>(), winners, counts);
        else if (type == SPK_F16)
            convwta_kernel<_Float16><<<grid, threads, 0, s>>>(
                (const _Float16 *)x, int32_t(t), int32_t(c), int32_t(h), int32_t(w), int32_t(ry), int32_t(rx),
                int32_t(count), _Float16(threshold), temp.as<float>(), winners, counts);
        else if (type == SPK_F64)
            convwta_kernel<double><<<grid, threads, 0, s>>>(
                (const double *)x, int32_t(t), int32_t(c), int32_t(h), int32_t(w), int32_t(ry), int32_t(rx),
                int32_t(count), threshold, temp.as<double>(), winners, counts);
        else if (type == SPK_U8)
            convwta_u8_kernel<<<dim3(unsigned(n)), threads, 0, s>>>(
                (const uint8_t *)x, int32_t(t), int32_t(c), int32_t(h), int32_t(w), int32_t(ry), int32_t(rx),
                int32_t(count), temp.as<float>(), winners, counts);
        else
            SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "ConvWTA: data type " << type_name(type) << " is not supported.");
        SPK_LAUNCHED();
    });
}

int32_t spk_rank_fcwta(const void *x, int32_t type, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                       double threshold, spk_winner *winners, int64_t *counts, void *stream)
{
    return guard([&] {
        SPK_REQUIRE(type == SPK_F16 || type == SPK_F32 || type == SPK_F64 || type == SPK_U8, SPK_ERR_UNSUPPORTED,
                    "FCWTA: data type " << type_name(type) << " is not supported.");
        SPK_ARG(radius >= 0, "Radius can't be less than 0.");
        SPK_ARG(count > 0, "Winners count can't be less than 1.");
        SPK_ARG(n >= 0 && t >= 1 && f >= 1 && f < INT32_MAX, "FCWTA: input shape is invalid.");
        if (n == 0) return;
        hipStream_t s = S(stream);
        Scratch temp(n * f * int64_t(type == SPK_F64 ? sizeof(double) : sizeof(float)), s);
        const dim3 grid = dim3(unsigned(n));
        const int threads = wta_threads(t * f);
        if (type == SPK_U8)
            fcwta_u8_kernel<<<grid, threads, 0, s>>>((const uint8_t *)x, int32_t(t), int32_t(f), int32_t(radius),
                                                      int32_t(count), temp.as<float>(), winners, counts);
        else if (type == SPK_F32)
            fcwta_kernel<float><<<grid, threads, 0, s>>>((const float *)x, int32_t(t), int32_t(f), int32_t(radius),
                                                          int32_t(count), float(threshold), temp.as<float>(), winners,
                                                          counts);
        else if (type == SPK_F16)
            fcwta_kernel<_Float16><<<grid, threads, 0, s>>>((const _Float16 *)x, int32_t(t), int32_t(f),
                                                             int32_t(radius), int32_t(count), _Float16(threshold),
                                                             temp.as<float>(), winners, counts);
        else
            fcwta_kernel<double><<<grid, threads, 0, s>>>((const double *)x, int32_t(t), int32_t(f), int32_t(radius),
                                                           int32_t(count), threshold, temp.as<double>(), winners,
                                                           counts);
        SPK_LAUNCHED();
    });
}
}
