// Rank STDP for conv and FC kernels (src/spyker/cpu/stdp.cpp:42-169).
//
// The reference walks samples in order and, inside a sample, winners in
// list order, updating kernel[winner.z] in place (stdp.cpp:139-154).  Each
// update of one weight depends only on that weight and on one bit: whether
// the input at its tap fired no later than the winner's output neuron
// (isum <= osum).  So the work splits in two launches:
//
//   records  one block per winner slot computes osum and, for every tap, the
//            bit isum <= osum, packed 32 to a word -- fully parallel;
//   apply    one block per output channel (conv) or output neuron (FC)
//            gathers the winners that name it, in (sample, rank) order, and
//            runs the sequential chain of updates with the weights held in
//            registers and the bits staged in LDS.
//
// Updates to different channels touch disjoint weights, so running the
// channels in parallel keeps the reference's order along every weight.
//
// The update is the reference's F64 formula (stdp.cpp:121-137), including
// the fused multiply-add its build contracts `old + rate*value` into, rounded
// once to the F32 weight.  Winners naming a configuration that does not
// exist fail the call before any weight changes (the reference throws at the
// first such winner, after applying the earlier ones).
#include "stdp.hpp"

namespace spk
{
namespace
{
constexpr int APPLY_THREADS = 1024;
constexpr int SLICE = APPLY_THREADS;  // weights per block, one per thread

// Packs one bit per lane of a 64-wide pass into the record words.
__device__ __forceinline__ void store_bits(uint32_t *rec, int32_t words, int32_t base, bool bit)
{
    const uint64_t ballot = __ballot(bit);
    if ((threadIdx.x & 63) == 0)
    {
        if (base / 32 < words) rec[base / 32] = uint32_t(ballot);
        if (base / 32 + 1 < words) rec[base / 32 + 1] = uint32_t(ballot >> 32);
    }
}

// One block (256 threads) per winner slot.
__global__ void __launch_bounds__(256)
    conv_records(const uint8_t *__restrict__ in, const uint8_t *__restrict__ out, StdpGeom g,
                 const spk_winner *__restrict__ win, const int64_t *__restrict__ counts,
                 uint32_t *__restrict__ bits, int2 *__restrict__ slot, int32_t *__restrict__ flag)
{
    const int64_t e = blockIdx.x, s = e / g.maxcount, k = e % g.maxcount;
    const spk_winner w = k < counts[s] ? win[e] : spk_winner{-1, 0, 0, 0, 0};
    int32_t bad = 0;
    if (w.c >= g.ncfg) bad = 1;
    else if (w.c >= 0 && (w.z < 0 || w.z >= g.cout || w.y < 0 || w.y >= g.ho || w.x < 0 || w.x >= g.wo)) bad = 2;
    if (threadIdx.x == 0)
    {
        slot[e] = make_int2(w.c >= 0 && !bad ? int32_t(w.z) : -1, int32_t(w.c));
        if (bad) atomicMax(flag, bad);
    }
    if (w.c < 0 || bad) return;
    const int64_t out_plane = int64_t(g.ho) * g.wo, in_plane = int64_t(g.h) * g.w, in_step = g.cin * in_plane;
    const uint8_t *os = out + (s * g.t * g.cout + w.z) * out_plane + w.y * g.wo + w.x;
    // first spiking step, scanned downwards as the reference does; the loads
    // are independent, so they are all in flight at once
    int32_t osum = g.t;
    for (int32_t i = g.t - 1; i >= 0; --i) osum = os[int64_t(i) * g.cout * out_plane] ? i : osum;
    const int32_t taps = g.kh * g.kw, ksz = g.cin * taps, words = (ksz + 31) / 32;
    const uint8_t *is = in + s * g.t * in_step;
    uint32_t *rec = bits + e * words;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int32_t base = wave * 64; base < words * 32; base += 256)
    {
        const int32_t i = base + lane;
        bool bit = false;
        if (i < ksz)
        {
            const int32_t ch = i / taps, a = (i / g.kw) % g.kh, b = i % g.kw;
            const int32_t iy = int32_t(w.y) * g.sy + a - g.pt, ix = int32_t(w.x) * g.sx + b - g.pl;
            int32_t isum = g.t;
            if (iy >= 0 && iy < g.h && ix >= 0 && ix < g.w)
            {
                const uint8_t *p = is + ch * in_plane + int64_t(iy) * g.w + ix;
                for (int32_t j = g.t - 1; j >= 0; --j) isum = p[int64_t(j) * in_step] ? j : isum;
            }
            bit = isum <= osum;
        }
        store_bits(rec, words, base, bit);
    }
}

__global__ void __launch_bounds__(256)
    fc_records(const uint8_t *__restrict__ in, const uint8_t *__restrict__ out, StdpGeom g,
               const spk_winner *__restrict__ win, const int64_t *__restrict__ counts, uint32_t *__restrict__ bits,
               int2 *__restrict__ slot, int32_t *__restrict__ flag)
{
    const int64_t e = blockIdx.x, s = e / g.maxcount, k = e % g.maxcount;
    const spk_winner w = k < counts[s] ? win[e] : spk_winner{-1, 0, 0, 0, 0};
    int32_t bad = 0;
    if (w.c >= g.ncfg) bad = 1;
    else if (w.c >= 0 && (w.x < 0 || w.x >= g.cout)) bad = 2;
    if (threadIdx.x == 0)
    {
        slot[e] = make_int2(w.c >= 0 && !bad ? int32_t(w.x) : -1, int32_t(w.c));
        if (bad) atomicMax(flag, bad);
    }
    if (w.c < 0 || bad) return;
    const uint8_t *os = out + s * g.t * int64_t(g.cout) + w.x;
    int32_t osum = g.t;
    for (int32_t i = g.t - 1; i >= 0; --i) osum = os[int64_t(i) * g.cout] ? i : osum;
    const int32_t fin = g.cin, words = (fin + 31) / 32;
    const uint8_t *is = in + s * g.t * int64_t(fin);
    uint32_t *rec = bits + e * words;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int32_t base = wave * 64; base < words * 32; base += 256)
    {
        const int32_t f = base + lane;
        bool bit = false;
        if (f < fin)
        {
            int32_t isum = g.t;
            for (int32_t j = g.t - 1; j >= 0; --j) isum = is[int64_t(j) * fin + f] ? j : isum;
            bit = isum <= osum;
        }
        store_bits(rec, words, base, bit);
    }
}

// One block per (output neuron, slice of SLICE weights): scans the slot
// keys written by the record kernel (winner.z for conv, winner.x for FC),
// stages the bits of the matching winners of each chunk in LDS, and applies
// them in slot order -- (sample, rank) order -- with each weight in a register.
template <typename K, bool SPARSE_RULE>
__global__ void __launch_bounds__(APPLY_THREADS)
    apply_records(K *__restrict__ kernel, StdpGeom g, StdpConfigs cfg, const int2 *__restrict__ slot,
                  const uint32_t *__restrict__ bits, const int32_t *__restrict__ flag, int32_t by_x)
{
    extern __shared__ uint32_t lds_bits[];  // [APPLY_THREADS][sw]
    __shared__ int32_t list[APPLY_THREADS];
    __shared__ int32_t cfgidx[APPLY_THREADS];
    __shared__ int32_t wave_count[APPLY_THREADS / 64];
    __shared__ spk_stdp_config lcfg[SPK_MAX_CONFIGS];
    if (*flag != 0) return;
    for (int32_t k = threadIdx.x; k < cfg.n; k += blockDim.x) lcfg[k] = cfg.c[k];  // read after the chunk barriers
    const int32_t key = blockIdx.x, w0 = blockIdx.y * SLICE;
    const int32_t ksz = by_x ? g.cin : g.cin * g.kh * g.kw, words = (ksz + 31) / 32;
    const int32_t nw = min(SLICE, ksz - w0), sw = (nw + 31) / 32, word0 = w0 / 32;
    K *kz = kernel + int64_t(key) * ksz + w0;
    const int32_t i = threadIdx.x;
    K wr = i < nw ? kz[i] : K(0);

    const int64_t total = g.n * g.maxcount;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t e0 = 0; e0 < total; e0 += APPLY_THREADS)
    {
        // order-preserving compaction of this chunk's slots that name `key`
        const int64_t e = e0 + threadIdx.x;
        const int2 sl = e < total ? slot[e] : make_int2(-1, 0);
        const bool match = sl.x == key;
        const uint64_t ballot = __ballot(match);
        if (lane == 0) wave_count[wave] = __popcll(ballot);
        __syncthreads();
        int32_t base = 0, m = 0;
        for (int v = 0; v < APPLY_THREADS / 64; ++v)
        {
            if (v < wave) base += wave_count[v];
            m += wave_count[v];
        }
        if (match)
        {
            const int32_t at = base + __popcll(ballot & ((1ull << lane) - 1));
            list[at] = int32_t(threadIdx.x);
            cfgidx[at] = sl.y;
        }
        __syncthreads();
        for (int32_t q = threadIdx.x; q < m * sw; q += APPLY_THREADS)
            lds_bits[q] = bits[(e0 + list[q / sw]) * words + word0 + q % sw];
        __syncthreads();
        if (i < nw && m > 0)
        {
            // the chain of updates on this weight, with the next winner's bit
            // and configuration fetched one update ahead
            uint32_t word = lds_bits[i >> 5];
            int32_t ci = cfgidx[0];
            for (int32_t q = 0; q < m; ++q)
            {
                const uint32_t cur = word;
                const spk_stdp_config cf = lcfg[ci];
                if (q + 1 < m) word = lds_bits[(q + 1) * sw + (i >> 5)], ci = cfgidx[q + 1];
                if constexpr (SPARSE_RULE)
                    wr = sparse_stdp_update(wr, (cur >> (i & 31)) & 1u, cf);
                else
                    wr = stdp_update<K>(wr, (cur >> (i & 31)) & 1u, cf);
            }
        }
        __syncthreads();
    }
    if (i < nw) kz[i] = wr;
}

// Reports a rejection found on the device (synchronises the stream).
void check_flag(int32_t *dflag, hipStream_t s)
{
    int32_t flag = 0;
    SPK_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SPK_HIP(hipStreamSynchronize(s));
    SPK_REQUIRE(flag != 1, SPK_ERR_CONFIG, "STDP configuration index is out of range.");
    SPK_REQUIRE(flag == 0, SPK_ERR_ARGUMENT, "Winner index is out of range.");
}

__global__ void clear_flag(int32_t *flag) { *flag = 0; }
}  // namespace

StdpConfigs pack_configs(const spk_stdp_config *configs, int64_t nconfigs)
{
    SPK_ARG(nconfigs > 0, "At least one STDP Config must be given.");
    SPK_ARG(nconfigs <= SPK_MAX_CONFIGS, "At most " << SPK_MAX_CONFIGS << " STDP configurations are supported.");
    StdpConfigs c{};
    for (int64_t i = 0; i < nconfigs; ++i) c.c[i] = configs[i];
    c.n = int32_t(nconfigs);
    return c;
}

// bits [slots][words] (u32, rounded to an even count) then slot keys [slots] (int2)
static int64_t bits_words(const StdpGeom &g, bool fc)
{
    const int64_t ksz = fc ? g.cin : int64_t(g.cin) * g.kh * g.kw;
    return (g.n * g.maxcount * ((ksz + 31) / 32) + 1) / 2 * 2;
}

int64_t stdp_bits_bytes(const StdpGeom &g, bool fc)
{
    return bits_words(g, fc) * int64_t(sizeof(uint32_t)) + g.n * g.maxcount * int64_t(sizeof(int2));
}

void stdp_launch(const uint8_t *in, void *kernel, int32_t ktype, const uint8_t *out, const StdpGeom &g,
                 const StdpConfigs &cfg, const spk_winner *winners, const int64_t *counts, uint32_t *bits,
                 int32_t *flag, bool fc, hipStream_t s)
{
    const int64_t ksz = fc ? g.cin : int64_t(g.cin) * g.kh * g.kw;
    SPK_ARG(ksz < (1LL << 30), "STDP: kernel too large.");
    const int64_t slots = g.n * g.maxcount;
    SPK_ARG(slots < (1LL << 31), "STDP: too many winner slots.");
    int2 *slot = reinterpret_cast<int2 *>(bits + bits_words(g, fc));
    if (fc)
        fc_records<<<dim3(unsigned(slots)), 256, 0, s>>>(in, out, g, winners, counts, bits, slot, flag);
    else
        conv_records<<<dim3(unsigned(slots)), 256, 0, s>>>(in, out, g, winners, counts, bits, slot, flag);
    SPK_LAUNCHED();
    stdp_apply(kernel, ktype, g, cfg, slot, bits, flag, fc, false, s);
}

template <typename K, bool SPARSE_RULE>
static void apply_launch(void *kernel, const StdpGeom &g, const StdpConfigs &cfg, const int2 *slot,
                         const uint32_t *bits, const int32_t *flag, bool fc, hipStream_t s)
{
    const int64_t ksz = fc ? g.cin : int64_t(g.cin) * g.kh * g.kw;
    const int64_t sw = (ksz < SLICE ? (ksz + 31) / 32 : SLICE / 32);
    const size_t lds = size_t(APPLY_THREADS) * sw * sizeof(uint32_t);
    if (lds > 65536)
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&apply_records<K, SPARSE_RULE>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    const dim3 grid(unsigned(g.cout), unsigned(cdiv(ksz, SLICE)));
    apply_records<K, SPARSE_RULE><<<grid, APPLY_THREADS, lds, s>>>((K *)kernel, g, cfg, slot, bits, flag, fc ? 1 : 0);
    SPK_LAUNCHED();
}

void stdp_apply(void *kernel, int32_t ktype, const StdpGeom &g, const StdpConfigs &cfg, const int2 *slot,
                const uint32_t *bits, const int32_t *flag, bool fc, bool sparse_rule, hipStream_t s)
{
    if (sparse_rule)
    {
        SPK_ARG(ktype == SPK_F32, "Sparse STDP: the kernel must be F32.");
        return apply_launch<float, true>(kernel, g, cfg, slot, bits, flag, fc, s);
    }
    if (ktype == SPK_F32) return apply_launch<float, false>(kernel, g, cfg, slot, bits, flag, fc, s);
    if (ktype == SPK_F16) return apply_launch<_Float16, false>(kernel, g, cfg, slot, bits, flag, fc, s);
    if (ktype == SPK_F64) return apply_launch<double, false>(kernel, g, cfg, slot, bits, flag, fc, s);
    SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "STDP: kernel data type " << type_name(ktype) << " is not supported.");
}
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
        SPK_REQUIRE(kernel_type == SPK_F16 || kernel_type == SPK_F32 || kernel_type == SPK_F64, SPK_ERR_UNSUPPORTED,
                    "STDP: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(n > 0, "Winners can't be empty.");
        SPK_ARG(t >= 1 && cin >= 1 && h >= 1 && w >= 1 && cout >= 1 && kh >= 1 && kw >= 1, "STDP: invalid shapes.");
        SPK_ARG(sy >= 1 && sx >= 1 && pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "STDP: invalid stride or pad.");
        SPK_ARG(ho == (h + pt + pb - kh) / sy + 1 && wo == (w + pl + pr - kw) / sx + 1,
                "Tensor shapes are not compatible.");
        SPK_ARG(maxcount >= 1, "STDP: winner capacity must be positive.");
        StdpConfigs c = pack_configs(configs, nconfigs);
        StdpGeom g{int32_t(t), int32_t(cin), int32_t(h), int32_t(w), int32_t(cout), int32_t(kh), int32_t(kw),
                   int32_t(ho), int32_t(wo), int32_t(sy), int32_t(sx), int32_t(pt), int32_t(pl), n, maxcount,
                   int32_t(nconfigs)};
        hipStream_t s = S(stream);
        Scratch flag(sizeof(int32_t), s), bits(stdp_bits_bytes(g, false), s);
        clear_flag<<<1, 1, 0, s>>>(flag.as<int32_t>());
        stdp_launch((const uint8_t *)in, kernel, kernel_type, (const uint8_t *)out, g, c, winners, counts,
                    bits.as<uint32_t>(), flag.as<int32_t>(), false, s);
        check_flag(flag.as<int32_t>(), s);
    });
}

int32_t spk_rank_fcstdp(const void *in, void *kernel, int32_t kernel_type, const void *out, int64_t n, int64_t t,
                        int64_t fin, int64_t fout, const spk_stdp_config *configs, int64_t nconfigs,
                        const spk_winner *winners, const int64_t *counts, int64_t maxcount, void *stream)
{
    return guard([&] {
        SPK_REQUIRE(kernel_type == SPK_F16 || kernel_type == SPK_F32 || kernel_type == SPK_F64, SPK_ERR_UNSUPPORTED,
                    "STDP: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(n > 0, "Winners can't be empty.");
        SPK_ARG(t >= 1 && fin >= 1 && fout >= 1, "STDP: invalid shapes.");
        SPK_ARG(maxcount >= 1, "STDP: winner capacity must be positive.");
        StdpConfigs c = pack_configs(configs, nconfigs);
        StdpGeom g{int32_t(t), int32_t(fin), 1, 1, int32_t(fout), 1, 1, 1, 1, 1, 1, 0, 0, n, maxcount,
                   int32_t(nconfigs)};
        hipStream_t s = S(stream);
        Scratch flag(sizeof(int32_t), s), bits(stdp_bits_bytes(g, true), s);
        clear_flag<<<1, 1, 0, s>>>(flag.as<int32_t>());
        stdp_launch((const uint8_t *)in, kernel, kernel_type, (const uint8_t *)out, g, c, winners, counts,
                    bits.as<uint32_t>(), flag.as<int32_t>(), true, s);
        check_flag(flag.as<int32_t>(), s);
    });
}
}
