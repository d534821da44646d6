// Direct cross-correlation for the plain path (F16, F32 and F64 kernels),
// plus DoG, Gabor and FC.
//
// Semantics: src/spyker/cpu/conv.cpp (NCHW cross-correlation, zero pad
// Len4{top,left,bottom,right}, stride), src/spyker/cpu/dog.cpp:39-96,
// src/spyker/cpu/fc.cpp.  As the reference does (opers.cpp:408-421), the
// input is first converted to the kernel's type and the output has the
// kernel's type.  The reference's arithmetic lives in oneDNN/cuDNN; this
// backend fixes one order for every output: acc = +0, then one fused
// multiply-add per tap in (cin, ky, kx) nesting order, in FP32 for F32 and
// F16 kernels (an F16 output is that FP32 sum rounded once) and in FP64 for
// F64 kernels.  Taps that fall in the zero pad are included: fma(0, w, acc)
// == acc for finite w because acc is never -0 (it starts at +0 and
// round-to-nearest turns exact cancellation into +0), so the result equals
// the tap-skipping form of the CPU oracle.
//
// Layout: a block owns a TH x TW tile of output pixels of one image and CT
// output channels, CT = 8 for layers of at most 8 output channels (filter
// banks: DoG, LoG, Gabor) and 32 otherwise.  The input window of the tile
// (with halo) is staged in LDS one channel chunk at a time; the weights of
// the CT channels are read through the scalar cache from a
// [cin][kh][kw][cout_pad] transposed copy, so each tap costs one LDS read and
// CT FMAs with SGPR operands.  When one block holds every filter of a DoG or
// LoG bank, the on/off combination runs in its epilogue (no intermediate
// planes).
#include "common.hpp"

namespace spk
{
namespace
{
constexpr int COT = 32;      // output channels per block
constexpr int COT_NARROW = 8; // ... for layers of at most 8 output channels
constexpr int THREADS = 256; // one output pixel per thread
constexpr int LDS_FLOATS = 8192;  // window elements a block stages at most

struct ConvGeom
{
    int32_t cin, h, w, cout, kh, kw, sy, sx, pt, pl, ho, wo;
    int32_t tw, th, rh, rw, cc;  // tile, window, channels per chunk
    int32_t tiles_x, tiles;
    int32_t cout_pad, cot;  // channels padded to a multiple of cot, the channels of a block
};

ConvGeom geometry(int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx,
                  int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    ConvGeom g{};
    g.cin = int32_t(cin), g.h = int32_t(h), g.w = int32_t(w), g.cout = int32_t(cout);
    g.kh = int32_t(kh), g.kw = int32_t(kw), g.sy = int32_t(sy), g.sx = int32_t(sx), g.pt = int32_t(pt),
    g.pl = int32_t(pl);
    g.ho = int32_t((h + pt + pb - kh) / sy + 1), g.wo = int32_t((w + pl + pr - kw) / sx + 1);
    g.tw = g.wo <= THREADS ? g.wo : 64;
    g.th = THREADS / g.tw;
    if (g.th > g.ho) g.th = g.ho;
    g.rh = (g.th - 1) * g.sy + g.kh, g.rw = (g.tw - 1) * g.sx + g.kw;
    SPK_ARG(g.rh * g.rw <= LDS_FLOATS, "Conv: kernel window too large for one tile.");
    g.cc = LDS_FLOATS / (g.rh * g.rw);
    if (g.cc > g.cin) g.cc = g.cin;
    g.tiles_x = int32_t(cdiv(g.wo, g.tw));
    g.tiles = g.tiles_x * int32_t(cdiv(g.ho, g.th));
    g.cot = g.cout <= COT_NARROW ? COT_NARROW : COT;
    g.cout_pad = int32_t(cdiv(g.cout, g.cot) * g.cot);
    return g;
}

// [cout][cin][kh][kw] -> [cin*kh*kw][cout_pad] (zero filled), widened to W
// (exact: F16 -> F32, F32 -> F32, F64 -> F64)
template <typename K, typename W>
__global__ void transpose_weights(const K *__restrict__ k, W *__restrict__ kt, int32_t cout, int32_t taps,
                                  int32_t cout_pad)
{
    const int64_t total = int64_t(taps) * cout_pad;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        int32_t o = int32_t(i % cout_pad), tap = int32_t(i / cout_pad);
        kt[i] = o < cout ? W(k[int64_t(o) * taps + tap]) : W(0);
    }
}

// An input element converted to the kernel's type K, held as the window type W.
template <typename K, typename W, typename I>
__device__ __forceinline__ W load_in(const I *p, int64_t i)
{
    return W(K(p[i]));
}

// The window type of a kernel type: F64 kernels keep F64 inputs.
template <typename K>
using WinT = typename std::conditional<std::is_same<K, double>::value, double, float>::type;

// The DoG (EPI_DOG) and LoG (EPI_LOG) on/off combination of one output pixel
// from the accumulators of a block holding all 2 * CNT filters: filter j
// against filter CNT + j, d = a - b, on = (d < 0 ? 0 : d) and for LoG off =
// (-d < 0 ? 0 : -d) (cpu/dog.cpp:62-96), out laid out [CNT (on) | CNT (off)]
// planes per image as the two-pass form writes them.
constexpr int EPI_NONE = 0, EPI_DOG = 1, EPI_LOG = 2;

template <int CNT, int EPI, int CT>
__device__ __forceinline__ void combine_pixel(const float (&acc)[CT], float *out, int64_t n, int64_t hw, int64_t pix)
{
    static_assert(2 * CNT <= CT, "the block must hold both filters of every pair");
    float *dst = out + n * (EPI == EPI_LOG ? 2 * CNT : CNT) * hw + pix;
#pragma unroll
    for (int j = 0; j < CNT; ++j)
    {
        const float d = acc[j] - acc[CNT + j], e = -d;
        dst[j * hw] = d < 0.0f ? 0.0f : d;
        if (EPI == EPI_LOG) dst[(CNT + j) * hw] = e < 0.0f ? 0.0f : e;
    }
}

// K: kernel type; Acc: float (F16 and F32 kernels) or double (F64 kernels
// and the FP64 check path of F32 kernels); O: output type; CT: output
// channels a block; EPI: plain output or a DoG/LoG combination (CT = 8, one
// block of channels, cout = 2 * count).
template <typename I, typename K, typename Acc, typename O, int CT, int EPI>
__global__ void __launch_bounds__(THREADS) conv_direct(const I *__restrict__ in, const WinT<K> *__restrict__ kt,
                                                       O *__restrict__ out, ConvGeom g)
{
    using W = WinT<K>;
    // the window of cc channels, sized at launch (a one-channel filter bank
    // stages a few KB, so more blocks fit a CU than a fixed array allows)
    extern __shared__ __align__(16) unsigned char conv_lds[];
    W *const win = reinterpret_cast<W *>(conv_lds);
    const int32_t tile = blockIdx.x, o0 = blockIdx.y * CT;
    const int64_t n = blockIdx.z;
    const int32_t ty0 = (tile / g.tiles_x) * g.th, tx0 = (tile % g.tiles_x) * g.tw;
    const int32_t ly = threadIdx.x / g.tw, lx = threadIdx.x % g.tw;
    const int32_t oy = ty0 + ly, ox = tx0 + lx;
    const bool active = threadIdx.x < g.th * g.tw && oy < g.ho && ox < g.wo;
    const int32_t iy0 = ty0 * g.sy - g.pt, ix0 = tx0 * g.sx - g.pl;
    const int32_t taps = g.kh * g.kw;

    Acc acc[CT];
#pragma unroll
    for (int j = 0; j < CT; ++j) acc[j] = Acc(0);

    const I *src = in + n * g.cin * int64_t(g.h) * g.w;
    for (int32_t c0 = 0; c0 < g.cin; c0 += g.cc)
    {
        const int32_t cc = min(g.cc, g.cin - c0);
        __syncthreads();
        const int32_t area = g.rh * g.rw;
        for (int32_t i = threadIdx.x; i < cc * area; i += THREADS)
        {
            const int32_t c = i / area, r = (i % area) / g.rw, q = i % g.rw;
            const int32_t iy = iy0 + r, ix = ix0 + q;
            win[i] = (iy >= 0 && iy < g.h && ix >= 0 && ix < g.w)
                         ? load_in<K, W>(src, (int64_t(c0 + c) * g.h + iy) * g.w + ix)
                         : W(0);
        }
        __syncthreads();
        if (active)
        {
            for (int32_t c = 0; c < cc; ++c)
            {
                const W *wrow = win + c * area + (ly * g.sy) * g.rw + lx * g.sx;
                const W *kc = kt + (int64_t(c0 + c) * taps) * g.cout_pad + o0;
                auto tap = [&](W x, const W *kk) {
#pragma unroll
                    for (int j = 0; j < CT; ++j)
                    {
                        if constexpr (sizeof(Acc) == 4)
                            acc[j] = __builtin_fmaf(x, kk[j], acc[j]);
                        else
                            acc[j] = __builtin_fma(double(x), double(kk[j]), acc[j]);
                    }
                };
                for (int32_t a = 0; a < g.kh; ++a)
                    for (int32_t b = 0; b < g.kw; ++b) tap(wrow[a * g.rw + b], kc + (a * g.kw + b) * g.cout_pad);
            }
        }
    }
    if (!active) return;
    if constexpr (EPI == EPI_NONE)
    {
        O *dst = out + (n * g.cout + o0) * int64_t(g.ho) * g.wo + int64_t(oy) * g.wo + ox;
#pragma unroll
        for (int j = 0; j < CT; ++j)
            if (o0 + j < g.cout) dst[int64_t(j) * g.ho * g.wo] = O(acc[j]);
    }
    else
    {
        static_assert(std::is_same<Acc, float>::value && std::is_same<O, float>::value, "F32 filter banks only");
        const int64_t hw = int64_t(g.ho) * g.wo, pix = int64_t(oy) * g.wo + ox;
        switch (g.cout / 2)
        {
        case 1: combine_pixel<1, EPI, CT>(acc, out, n, hw, pix); break;
        case 2: combine_pixel<2, EPI, CT>(acc, out, n, hw, pix); break;
        case 3: combine_pixel<3, EPI, CT>(acc, out, n, hw, pix); break;
        default: combine_pixel<4, EPI, CT>(acc, out, n, hw, pix); break;
        }
    }
}

// DoG combine: mid is [planes][2][count*ho*wo]; out = max(g1 - g2, 0) as
// std::max<T>(d, 0) evaluates it (d < 0 ? 0 : d keeps -0 and NaN).
__global__ void dog_combine(const float *__restrict__ mid, float *__restrict__ out, int64_t planes, int64_t len)
{
    const int64_t total = planes * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t p = i / len, j = i % len;
        const float d = mid[(p * 2) * len + j] - mid[(p * 2 + 1) * len + j];
        out[i] = d < 0.0f ? 0.0f : d;
    }
}

// LoG on/off pair (cpu/dog.cpp:62-84): with d = conv(std/sqrt2) - conv(std*sqrt2),
// on = max(d, 0) and off = max(-d, 0) in std::max's form (a < b ? b : a, so a
// zero difference keeps its sign).  mid and out are [planes][2][len].
__global__ void log_combine(const float *__restrict__ mid, float *__restrict__ out, int64_t planes, int64_t len)
{
    const int64_t total = planes * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t p = i / len, j = i % len;
        const float d = mid[(p * 2) * len + j] - mid[(p * 2 + 1) * len + j], e = -d;
        out[(p * 2) * len + j] = d < 0.0f ? 0.0f : d;
        out[(p * 2 + 1) * len + j] = e < 0.0f ? 0.0f : e;
    }
}

// FC: kt is [fin][fout]; out[r][o] = fma chain over f in order.
template <typename I>
__global__ void fc_kernel(const I *__restrict__ in, const float *__restrict__ kt, float *__restrict__ out,
                          int64_t rows, int32_t fin, int32_t fout)
{
    const int64_t r = blockIdx.y;
    const int32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= fout) return;
    const I *row = in + r * fin;
    float acc = 0.0f;
    for (int32_t f = 0; f < fin; ++f) acc = __builtin_fmaf(float(row[f]), kt[int64_t(f) * fout + o], acc);
    out[r * fout + o] = acc;
}

__global__ void transpose_fc(const float *__restrict__ k, float *__restrict__ kt, int32_t fout, int32_t fin)
{
    const int64_t total = int64_t(fout) * fin;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        int32_t o = int32_t(i % fout), f = int32_t(i / fout);
        kt[i] = k[int64_t(o) * fin + f];
    }
}

template <typename I, typename K, typename Acc, typename O, int EPI = EPI_NONE>
void conv_run(const void *in, const void *kernel, void *out, int64_t n, const ConvGeom &g, hipStream_t s)
{
    using W = WinT<K>;
    const int32_t taps = g.cin * g.kh * g.kw;
    // the transposed weights live in the stream's workspace (calls on a stream
    // run in order), so a captured conv records no allocation
    W *kt = static_cast<W *>(stream_workspace(1, s, int64_t(taps) * g.cout_pad * int64_t(sizeof(W))));
    transpose_weights<K, W><<<flat_grid(int64_t(taps) * g.cout_pad, 256), 256, 0, s>>>((const K *)kernel, kt, g.cout,
                                                                                        taps, g.cout_pad);
    SPK_LAUNCHED();
    for (int64_t n0 = 0; n0 < n; n0 += 65535)
    {
        const int64_t nb = n - n0 < 65535 ? n - n0 : 65535;
        dim3 grid(unsigned(g.tiles), unsigned(g.cout_pad / g.cot), unsigned(nb));
        const I *src = (const I *)in + n0 * g.cin * int64_t(g.h) * g.w;
        O *dst = (O *)out + n0 * (EPI == EPI_DOG ? g.cout / 2 : g.cout) * int64_t(g.ho) * g.wo;
        const size_t lds = size_t(g.cc) * g.rh * g.rw * sizeof(W);
        if constexpr (EPI != EPI_NONE)
            conv_direct<I, K, Acc, O, COT_NARROW, EPI><<<grid, THREADS, lds, s>>>(src, kt, dst, g);
        else if (g.cot == COT_NARROW)
            conv_direct<I, K, Acc, O, COT_NARROW, EPI_NONE><<<grid, THREADS, lds, s>>>(src, kt, dst, g);
        else
            conv_direct<I, K, Acc, O, COT, EPI_NONE><<<grid, THREADS, lds, s>>>(src, kt, dst, g);
        SPK_LAUNCHED();
    }
}

void check_conv(int64_t n, int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw, int64_t sy,
                int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    SPK_ARG(sy >= 1 && sx >= 1, "Stride items are incorrect.");
    SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
    SPK_ARG(n >= 0 && cin >= 1 && h >= 1 && w >= 1 && cout >= 1 && kh >= 1 && kw >= 1,
            "Conv: tensor shapes are invalid.");
    SPK_ARG((h + pt + pb - kh) / sy + 1 > 0 && (w + pl + pr - kw) / sx + 1 > 0 && h + pt + pb >= kh &&
                w + pl + pr >= kw,
            "Input and kernel shapes are not compatible.");
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_conv(const void *in, int32_t in_type, const void *kernel, int32_t kernel_type, void *out,
                 int32_t out_type, int64_t n, int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw,
                 int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        check_conv(n, cin, h, w, cout, kh, kw, sy, sx, pt, pl, pb, pr);
        SPK_REQUIRE(kernel_type == SPK_F16 || kernel_type == SPK_F32 || kernel_type == SPK_F64, SPK_ERR_UNSUPPORTED,
                    "Conv: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_REQUIRE(out_type == kernel_type || (kernel_type == SPK_F32 && out_type == SPK_F64), SPK_ERR_UNSUPPORTED,
                    "Kernel and output must have the same type.");
        if (n == 0) return;
        ConvGeom g = geometry(cin, h, w, cout, kh, kw, sy, sx, pt, pl, pb, pr);
        hipStream_t s = S(stream);
        using H = _Float16;
        if (kernel_type == SPK_F32)
        {
            if (in_type == SPK_U8 && out_type == SPK_F32) return conv_run<uint8_t, float, float, float>(in, kernel, out, n, g, s);
            if (in_type == SPK_F32 && out_type == SPK_F32) return conv_run<float, float, float, float>(in, kernel, out, n, g, s);
            if (in_type == SPK_F16 && out_type == SPK_F32) return conv_run<H, float, float, float>(in, kernel, out, n, g, s);
            if (in_type == SPK_F64 && out_type == SPK_F32) return conv_run<double, float, float, float>(in, kernel, out, n, g, s);
            // the FP64 check path: F32 weights and inputs, FP64 sums
            if (in_type == SPK_U8 && out_type == SPK_F64) return conv_run<uint8_t, float, double, double>(in, kernel, out, n, g, s);
            if (in_type == SPK_F32 && out_type == SPK_F64) return conv_run<float, float, double, double>(in, kernel, out, n, g, s);
        }
        else if (kernel_type == SPK_F16)
        {
            if (in_type == SPK_U8) return conv_run<uint8_t, H, float, H>(in, kernel, out, n, g, s);
            if (in_type == SPK_F16) return conv_run<H, H, float, H>(in, kernel, out, n, g, s);
            if (in_type == SPK_F32) return conv_run<float, H, float, H>(in, kernel, out, n, g, s);
        }
        else
        {
            if (in_type == SPK_U8) return conv_run<uint8_t, double, double, double>(in, kernel, out, n, g, s);
            if (in_type == SPK_F32) return conv_run<float, double, double, double>(in, kernel, out, n, g, s);
            if (in_type == SPK_F64) return conv_run<double, double, double, double>(in, kernel, out, n, g, s);
        }
        SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED,
                    "Conv: data types " << type_name(in_type) << " -> " << type_name(out_type) << " are not supported.");
    });
}

int32_t spk_fc(const void *in, int32_t in_type, const void *kernel, void *out, int64_t rows, int64_t fin,
               int64_t fout, void *stream)
{
    return guard([&] {
        SPK_ARG(rows >= 0 && fin >= 1 && fout >= 1, "FC: tensor shapes are invalid.");
        SPK_ARG(rows < 65536LL * 65536LL, "FC: too many rows.");
        if (rows == 0) return;
        hipStream_t s = S(stream);
        Scratch kt(fin * fout * int64_t(sizeof(float)), s);
        transpose_fc<<<flat_grid(fin * fout, 256), 256, 0, s>>>((const float *)kernel, kt.as<float>(), int32_t(fout),
                                                                int32_t(fin));
        SPK_LAUNCHED();
        for (int64_t r0 = 0; r0 < rows; r0 += 65535)
        {
            const int64_t rb = rows - r0 < 65535 ? rows - r0 : 65535;
            dim3 grid(unsigned(cdiv(fout, 256)), unsigned(rb));
            if (in_type == SPK_U8)
                fc_kernel<uint8_t><<<grid, 256, 0, s>>>((const uint8_t *)in + r0 * fin, kt.as<float>(),
                                                        (float *)out + r0 * fout, rb, int32_t(fin), int32_t(fout));
            else if (in_type == SPK_F32)
                fc_kernel<float><<<grid, 256, 0, s>>>((const float *)in + r0 * fin, kt.as<float>(),
                                                      (float *)out + r0 * fout, rb, int32_t(fin), int32_t(fout));
            else
                SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "FC: data type " << type_name(in_type) << " is not supported.");
            SPK_LAUNCHED();
        }
    });
}

int32_t spk_dog(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        check_conv(planes, 1, h, w, 2 * count, k, k, 1, 1, pt, pl, pb, pr);
        if (planes == 0) return;
        ConvGeom g = geometry(1, h, w, 2 * count, k, k, 1, 1, pt, pl, pb, pr);
        hipStream_t s = S(stream);
        // one block holds every filter: the combination in the conv's epilogue
        if (g.cot == COT_NARROW) return conv_run<float, float, float, float, EPI_DOG>(in, kernel, out, planes, g, s);
        const int64_t len = count * int64_t(g.ho) * g.wo;
        Scratch mid(planes * 2 * len * int64_t(sizeof(float)), s);
        conv_run<float, float, float, float>(in, kernel, mid.ptr, planes, g, s);
        dog_combine<<<flat_grid(planes * len, 256), 256, 0, s>>>(mid.as<float>(), (float *)out, planes, len);
        SPK_LAUNCHED();
    });
}

int32_t spk_log(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        check_conv(planes, 1, h, w, 2 * count, k, k, 1, 1, pt, pl, pb, pr);
        if (planes == 0) return;
        ConvGeom g = geometry(1, h, w, 2 * count, k, k, 1, 1, pt, pl, pb, pr);
        hipStream_t s = S(stream);
        if (g.cot == COT_NARROW) return conv_run<float, float, float, float, EPI_LOG>(in, kernel, out, planes, g, s);
        const int64_t len = count * int64_t(g.ho) * g.wo;
        Scratch mid(planes * 2 * len * int64_t(sizeof(float)), s);
        conv_run<float, float, float, float>(in, kernel, mid.ptr, planes, g, s);
        log_combine<<<flat_grid(planes * len, 256), 256, 0, s>>>(mid.as<float>(), (float *)out, planes, len);
        SPK_LAUNCHED();
    });
}

int32_t spk_gabor(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                  int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        check_conv(planes, 1, h, w, count, k, k, 1, 1, pt, pl, pb, pr);
        if (planes == 0) return;
        ConvGeom g = geometry(1, h, w, count, k, k, 1, 1, pt, pl, pb, pr);
        conv_run<float, float, float, float>(in, kernel, out, planes, g, S(stream));
    });
}
}
