// Rank STDP for conv and FC kernels (src/spyker/cpu/stdp.cpp:42-169).
//
// The reference walks samples in order and, inside a sample, winners in
// list order, updating kernel[winner.z] in place (stdp.cpp:139-154).  Updates
// to different output channels touch disjoint weights, so only the order
// along one channel matters: one block owns one output channel (conv) or one
// output neuron (FC), compacts the winners that name it in (sample, rank)
// order, and applies them one after another.  Inside a block each thread owns
// a fixed set of weights, so consecutive updates need no barrier.
//
// The update itself is the reference's F64 formula (stdp.cpp:121-137),
// including the fused multiply-add its build contracts `old + rate*value`
// into, rounded once to the F32 master weight.
#include "common.hpp"

namespace spk
{
namespace
{
constexpr int MAXCFG = SPK_MAX_CONFIGS;

struct Configs
{
    spk_stdp_config c[MAXCFG];
    int32_t n;
};

__device__ __forceinline__ float stdp_update(float w, bool potentiate, const spk_stdp_config &cfg)
{
    const double old = double(w);
    double value = cfg.stabilize ? (old - cfg.lower) * (cfg.upper - old) : 1.0;
    value = __builtin_fma(potentiate ? cfg.positive : cfg.negative, value, old);
    const double hi = value < cfg.upper ? value : cfg.upper;
    const double r = cfg.lower < hi ? hi : cfg.lower;
    return float(r);
}

// Fails the whole call (no weight changes) if any active winner names a
// configuration that does not exist (flag 1) or a neuron outside the output
// (flag 2; the reference does not check this and would index out of range).
__device__ bool winners_valid(const spk_winner *__restrict__ win, const int64_t *__restrict__ counts, int64_t n,
                              int64_t maxcount, int32_t ncfg, int64_t zlim, int64_t ylim, int64_t xlim,
                              int32_t *flag)
{
    __shared__ int32_t bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int64_t e = threadIdx.x; e < n * maxcount; e += blockDim.x)
    {
        const int64_t s = e / maxcount, k = e % maxcount;
        if (k >= counts[s]) continue;
        const spk_winner w = win[e];
        if (w.c >= ncfg) atomicMax(&bad, 1);
        if (w.c >= 0 && (w.z < 0 || w.z >= zlim || w.y < 0 || w.y >= ylim || w.x < 0 || w.x >= xlim))
            atomicMax(&bad, 2);
    }
    __syncthreads();
    const int32_t b = bad;
    if (b != 0 && blockIdx.x == 0 && threadIdx.x == 0) *flag = b;
    return b == 0;
}

// Order-preserving compaction of one chunk of winner slots whose key equals
// `key`; returns the number of entries written to list.
__device__ int32_t compact(const spk_winner *__restrict__ win, const int64_t *__restrict__ counts, int64_t maxcount,
                           int64_t e0, int64_t total, int64_t key, bool by_x, int32_t *list, int32_t *wave_count)
{
    const int64_t e = e0 + threadIdx.x;
    bool match = false;
    if (e < total)
    {
        const int64_t s = e / maxcount, k = e % maxcount;
        if (k < counts[s])
        {
            const spk_winner w = win[e];
            match = w.c >= 0 && (by_x ? w.x : w.z) == key;
        }
    }
    const uint64_t ballot = __ballot(match);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_count[wave] = __popcll(ballot);
    __syncthreads();
    int32_t base = 0, total_n = 0;
    for (int i = 0; i < int(blockDim.x >> 6); ++i)
    {
        if (i < wave) base += wave_count[i];
        total_n += wave_count[i];
    }
    if (match) list[base + __popcll(ballot & ((1ull << lane) - 1))] = int32_t(threadIdx.x);
    __syncthreads();
    return total_n;
}

__global__ void __launch_bounds__(256)
    convstdp_kernel(const uint8_t *__restrict__ in, float *__restrict__ kernel, const uint8_t *__restrict__ out,
                    int32_t t, int32_t cin, int32_t h, int32_t w, int32_t cout, int32_t kh, int32_t kw, int32_t ho,
                    int32_t wo, Configs cfg, const spk_winner *__restrict__ win, const int64_t *__restrict__ counts,
                    int64_t n, int64_t maxcount, int32_t sy, int32_t sx, int32_t pt, int32_t pl,
                    int32_t *__restrict__ flag)
{
    __shared__ int32_t list[256];
    __shared__ int32_t wave_count[4];
    if (!winners_valid(win, counts, n, maxcount, cfg.n, cout, ho, wo, flag)) return;
    const int32_t z = blockIdx.x;
    const int32_t ksz = cin * kh * kw;
    float *kz = kernel + int64_t(z) * ksz;
    const int64_t total = n * maxcount;
    const int64_t in_plane = int64_t(h) * w, in_step = cin * in_plane, out_plane = int64_t(ho) * wo;
    for (int64_t e0 = 0; e0 < total; e0 += 256)
    {
        const int32_t m = compact(win, counts, maxcount, e0, total, z, false, list, wave_count);
        for (int32_t q = 0; q < m; ++q)
        {
            const int64_t e = e0 + list[q], s = e / maxcount;
            const spk_winner wv = win[e];
            const uint8_t *os = out + ((s * t) * cout + z) * out_plane + wv.y * wo + wv.x;
            int32_t osum = t;
            for (int32_t i = 0; i < t; ++i)
                if (os[int64_t(i) * cout * out_plane])
                {
                    osum = i;
                    break;
                }
            const spk_stdp_config &c = cfg.c[wv.c];
            const uint8_t *is = in + s * t * in_step;
            for (int32_t i = threadIdx.x; i < ksz; i += blockDim.x)
            {
                const int32_t ch = i / (kh * kw), a = (i / kw) % kh, b = i % kw;
                const int32_t iy = int32_t(wv.y) * sy + a - pt, ix = int32_t(wv.x) * sx + b - pl;
                int32_t isum = t;
                if (iy >= 0 && iy < h && ix >= 0 && ix < w)
                {
                    const uint8_t *p = is + ch * in_plane + int64_t(iy) * w + ix;
                    for (int32_t j = 0; j < t; ++j)
                        if (p[int64_t(j) * in_step])
                        {
                            isum = j;
                            break;
                        }
                }
                kz[i] = stdp_update(kz[i], isum <= osum, c);
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
    fcstdp_kernel(const uint8_t *__restrict__ in, float *__restrict__ kernel, const uint8_t *__restrict__ out,
                  int32_t t, int32_t fin, int32_t fout, Configs cfg, const spk_winner *__restrict__ win,
                  const int64_t *__restrict__ counts, int64_t n, int64_t maxcount, int32_t *__restrict__ flag)
{
    __shared__ int32_t list[256];
    __shared__ int32_t wave_count[4];
    if (!winners_valid(win, counts, n, maxcount, cfg.n, INT64_MAX, INT64_MAX, fout, flag)) return;
    const int32_t o = blockIdx.x;
    float *ko = kernel + int64_t(o) * fin;
    const int64_t total = n * maxcount;
    for (int64_t e0 = 0; e0 < total; e0 += 256)
    {
        const int32_t m = compact(win, counts, maxcount, e0, total, o, true, list, wave_count);
        for (int32_t q = 0; q < m; ++q)
        {
            const int64_t e = e0 + list[q], s = e / maxcount;
            const spk_winner wv = win[e];
            const uint8_t *os = out + s * t * int64_t(fout) + o;
            int32_t osum = t;
            for (int32_t i = 0; i < t; ++i)
                if (os[int64_t(i) * fout])
                {
                    osum = i;
                    break;
                }
            const spk_stdp_config &c = cfg.c[wv.c];
            const uint8_t *is = in + s * t * int64_t(fin);
            for (int32_t f = threadIdx.x; f < fin; f += blockDim.x)
            {
                int32_t isum = t;
                for (int32_t j = 0; j < t; ++j)
                    if (is[int64_t(j) * fin + f])
                    {
                        isum = j;
                        break;
                    }
                ko[f] = stdp_update(ko[f], isum <= osum, c);
            }
        }
        __syncthreads();
    }
}

Configs pack(const spk_stdp_config *configs, int64_t nconfigs)
{
    SPK_ARG(nconfigs > 0, "At least one STDP Config must be given.");
    SPK_ARG(nconfigs <= MAXCFG, "At most " << MAXCFG << " STDP configurations are supported.");
    Configs c{};
    for (int64_t i = 0; i < nconfigs; ++i) c.c[i] = configs[i];
    c.n = int32_t(nconfigs);
    return c;
}

// Reports a configuration error found on the device (synchronises the stream).
void check_flag(int32_t *dflag, hipStream_t s)
{
    int32_t flag = 0;
    SPK_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SPK_HIP(hipStreamSynchronize(s));
    SPK_REQUIRE(flag != 1, SPK_ERR_CONFIG, "STDP configuration index is out of range.");
    SPK_REQUIRE(flag == 0, SPK_ERR_ARGUMENT, "Winner index is out of range.");
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_rank_convstdp(const void *in, void *kernel, int32_t kernel_type, const void *out, int64_t n, int64_t t,
                          int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw, int64_t ho,
                          int64_t wo, const spk_stdp_config *configs, int64_t nconfigs, const spk_winner *winners,
                          const int64_t *counts, int64_t maxcount, int64_t sy, int64_t sx, int64_t pt, int64_t pl,
                          int64_t pb, int64_t pr, void *stream)
{
    return guard([&] {
        SPK_REQUIRE(kernel_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "STDP: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(n > 0, "Winners can't be empty.");
        SPK_ARG(t >= 1 && cin >= 1 && h >= 1 && w >= 1 && cout >= 1 && kh >= 1 && kw >= 1, "STDP: invalid shapes.");
        SPK_ARG(sy >= 1 && sx >= 1 && pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "STDP: invalid stride or pad.");
        SPK_ARG(ho == (h + pt + pb - kh) / sy + 1 && wo == (w + pl + pr - kw) / sx + 1,
                "Tensor shapes are not compatible.");
        SPK_ARG(maxcount >= 1, "STDP: winner capacity must be positive.");
        Configs c = pack(configs, nconfigs);
        hipStream_t s = S(stream);
        Scratch flag(sizeof(int32_t), s);
        SPK_HIP(hipMemsetAsync(flag.ptr, 0, sizeof(int32_t), s));
        convstdp_kernel<<<dim3(unsigned(cout)), 256, 0, s>>>(
            (const uint8_t *)in, (float *)kernel, (const uint8_t *)out, int32_t(t), int32_t(cin), int32_t(h),
            int32_t(w), int32_t(cout), int32_t(kh), int32_t(kw), int32_t(ho), int32_t(wo), c, winners, counts, n,
            maxcount, int32_t(sy), int32_t(sx), int32_t(pt), int32_t(pl), flag.as<int32_t>());
        SPK_LAUNCHED();
        check_flag(flag.as<int32_t>(), s);
    });
}

int32_t spk_rank_fcstdp(const void *in, void *kernel, int32_t kernel_type, const void *out, int64_t n, int64_t t,
                        int64_t fin, int64_t fout, const spk_stdp_config *configs, int64_t nconfigs,
                        const spk_winner *winners, const int64_t *counts, int64_t maxcount, void *stream)
{
    return guard([&] {
        SPK_REQUIRE(kernel_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "STDP: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(n > 0, "Winners can't be empty.");
        SPK_ARG(t >= 1 && fin >= 1 && fout >= 1, "STDP: invalid shapes.");
        SPK_ARG(maxcount >= 1, "STDP: winner capacity must be positive.");
        Configs c = pack(configs, nconfigs);
        hipStream_t s = S(stream);
        Scratch flag(sizeof(int32_t), s);
        SPK_HIP(hipMemsetAsync(flag.ptr, 0, sizeof(int32_t), s));
        fcstdp_kernel<<<dim3(unsigned(fout)), 256, 0, s>>>((const uint8_t *)in, (float *)kernel,
                                                           (const uint8_t *)out, int32_t(t), int32_t(fin),
                                                           int32_t(fout), c, winners, counts, n, maxcount,
                                                           flag.as<int32_t>());
        SPK_LAUNCHED();
        check_flag(flag.as<int32_t>(), s);
    });
}
}
