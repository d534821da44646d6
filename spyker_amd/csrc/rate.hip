// Rate (Poisson) coding, the reference's CPU algorithm run on the device
// (src/spyker/cpu/code.cpp:134-243).
//
// Each element first gets its rank-code spike time (latency or sorted, the
// same U16 times the rank code uses); that time picks a row of the Poisson
// interval table (row r holds intervals of mean r + 1).  Sample i draws from
// its own std::mt19937 seeded with rnd + i, rnd one draw of the shared
// generator per call: every element draws its first spike time in element
// order, then at each step t, in element order, every element whose next time
// is t counts a spike and draws its next interval.  The output at step t is
// the count so far.
//
// One workgroup codes one sample.  The draw order is sequential by
// definition, so the workgroup reproduces it in parallel: the elements that
// draw at step t are ranked by an ordered scan (wave ballots, then the wave
// totals), and the k-th of them takes stream word pos + k.  The MT19937
// stream itself is generated 624 words at a time by a four-phase parallel
// twist in LDS, tempered into a window that always holds the next `len`
// words.
//
// The reference's table has T rows while a spike time can be T (the sample
// minimum under latency coding) or more (values at the minimum under sorted
// coding), which reads past its table.  Here the table has T + 1 rows and
// times at or past T use row T (intervals of mean T + 1).
#include "code.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace spk
{
namespace
{
constexpr int RATE_THREADS = 256;
constexpr int MT_N = 624;
constexpr int POISSON_BINS = 1024;  // code.cpp:134
constexpr int64_t RATE_MAX_LEN = 16384;
constexpr int64_t RATE_MAX_TIME = 4096;  // the host table costs O(T^2 sqrt T)

__device__ __forceinline__ uint32_t twist_word(uint32_t a, uint32_t b, uint32_t c)
{
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ uint32_t temper(uint32_t z)
{
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    return z ^ (z >> 18);
}

// One MT19937 state transition over the 624 words in LDS, as the sequential
// loop would do it: words 0..226 read only old words; 227..453 read the new
// word 227 places back; 454..622 likewise; 623 reads the new words 0 and 396.
__device__ void mt_twist(uint32_t *mt)
{
    const int t = threadIdx.x;
    uint32_t v = 0;
    if (t < 227) v = twist_word(mt[t], mt[t + 1], mt[t + 397]);
    __syncthreads();
    if (t < 227) mt[t] = v;
    __syncthreads();
    if (t < 227) v = twist_word(mt[227 + t], mt[228 + t], mt[t]);
    __syncthreads();
    if (t < 227) mt[227 + t] = v;
    __syncthreads();
    if (t < 169) v = twist_word(mt[454 + t], mt[455 + t], mt[227 + t]);
    __syncthreads();
    if (t < 169) mt[454 + t] = v;
    __syncthreads();
    if (t == 0) mt[623] = twist_word(mt[623], mt[0], mt[396]);
    __syncthreads();
}

// Makes stream words [pos, need) readable at buf[k - base].  base and filled
// (one past the last generated word) are the same in every thread.
__device__ void ensure(uint32_t *mt, uint32_t *buf, int64_t cap, int64_t pos, int64_t need, int64_t &base,
                       int64_t &filled)
{
    while (filled < need)
    {
        if (filled - base + MT_N > cap)
        {
            // drop the consumed words [base, pos): move [pos, filled) to the front
            const int64_t keep = filled - pos, shift = pos - base;
            for (int64_t c = 0; c < keep; c += RATE_THREADS)
            {
                const int64_t k = c + threadIdx.x;
                uint32_t w = 0;
                if (k < keep) w = buf[shift + k];
                __syncthreads();
                if (k < keep) buf[k] = w;
                __syncthreads();
            }
            base = pos;
        }
        mt_twist(mt);
        for (int k = threadIdx.x; k < MT_N; k += RATE_THREADS) buf[filled - base + k] = temper(mt[k]);
        __syncthreads();
        filled += MT_N;
    }
}

// Exclusive prefix of v over the workgroup's threads in thread order, and its total.
__device__ __forceinline__ int32_t block_exclusive(int32_t v, int32_t *wsum, int32_t &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t incl = v;
    for (int d = 1; d < 64; d <<= 1)
    {
        const int32_t u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int32_t before = 0;
    total = 0;
    for (int k = 0; k < RATE_THREADS / 64; ++k)
    {
        before += k < wave ? wsum[k] : 0;
        total += wsum[k];
    }
    return before + incl - v;
}

// Thread k owns the contiguous elements [k * per, (k + 1) * per), so the
// draws of a step are ranked by one scan of per-thread counts and each thread
// issues its table reads back to back.  ROWS_IN_LDS keeps the rows (spike
// times) in LDS, when they fit, instead of re-reading them from HBM.
template <typename O, bool ROWS_IN_LDS>
__global__ void __launch_bounds__(RATE_THREADS)
    rate_code_kernel(const uint16_t *__restrict__ rows_g, const uint16_t *__restrict__ table, O *__restrict__ out,
                     int64_t len, int32_t t, uint64_t rnd)
{
    extern __shared__ uint32_t lds[];
    __shared__ int32_t wsum[RATE_THREADS / 64];
    const int64_t cap = len + MT_N;
    uint32_t *mt = lds, *buf = lds + MT_N;
    uint16_t *times = reinterpret_cast<uint16_t *>(buf + cap), *values = times + len, *rows_l = values + len;
    const int64_t s = blockIdx.x;
    const uint16_t *rows = rows_g + s * len;
    O *o = out + s * int64_t(t) * len;
    const uint32_t last = uint32_t(t);  // the table's last row
    if constexpr (ROWS_IN_LDS)
        for (int64_t j = threadIdx.x; j < len; j += RATE_THREADS) rows_l[j] = rows[j];
    // table position of element j's next interval for stream word `word`
    auto slot = [&](int64_t j, uint32_t word) {
        uint32_t r;
        if constexpr (ROWS_IN_LDS)
            r = rows_l[j];
        else
            r = rows[j];
        r = r < last ? r : last;
        return r * uint32_t(POISSON_BINS) + (word & (POISSON_BINS - 1));
    };
    const int64_t per = (len + RATE_THREADS - 1) / RATE_THREADS;
    const int64_t j0 = int64_t(threadIdx.x) * per < len ? int64_t(threadIdx.x) * per : len;
    const int64_t j1 = j0 + per < len ? j0 + per : len;

    if (threadIdx.x == 0)
    {
        // std::mt19937(seed): seed mod 2^32 (code.cpp:215 `Random random(rnd + i)`)
        uint32_t x = uint32_t(rnd + uint64_t(s));
        mt[0] = x;
        for (uint32_t i = 1; i < MT_N; ++i)
        {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            mt[i] = x;
        }
    }
    __syncthreads();

    int64_t base = 0, filled = 0, pos = 0;
    ensure(mt, buf, cap, 0, len, base, filled);
    for (int64_t j = j0; j < j1; ++j)
    {
        values[j] = 0;
        times[j] = table[slot(j, buf[j - base])];
    }
    pos = len;
    __syncthreads();

    for (int32_t step = 0; step < t; ++step)
    {
        ensure(mt, buf, cap, pos, pos + len, base, filled);
        const uint16_t now = uint16_t(step);
        int32_t cnt = 0;
        for (int64_t j = j0; j < j1; ++j) cnt += times[j] == now ? 1 : 0;
        int32_t total;
        int64_t at = pos + block_exclusive(cnt, wsum, total) - base;
        // eight elements at a time: stream words and table positions first, then
        // the eight table reads together, then the updates
        for (int64_t c = j0; c < j1; c += 8)
        {
            uint32_t idx[8];
            bool fire[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                const int64_t j = c + k;
                fire[k] = j < j1 && times[j] == now;
                idx[k] = fire[k] ? slot(j, buf[at]) : 0u;
                at += fire[k] ? 1 : 0;
            }
            uint16_t gap[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) gap[k] = fire[k] ? table[idx[k]] : uint16_t(0);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (fire[k])
                {
                    values[c + k] = uint16_t(values[c + k] + 1);
                    times[c + k] = uint16_t(times[c + k] + gap[k]);
                }
        }
        __syncthreads();
        for (int64_t j = threadIdx.x; j < len; j += RATE_THREADS) o[int64_t(step) * len + j] = O(values[j]);
        pos += total;
    }
}

// Device Poisson tables, one per (device, T), kept until spk_poisson_clear
// (the reference caches them per T the same way, code.cpp:156-166).
struct Tables
{
    std::mutex mutex;
    std::map<std::pair<int, int64_t>, uint16_t *> tables;
};

Tables &tables()
{
    static Tables *t = new Tables;
    return *t;
}

const uint16_t *device_table(int64_t time)
{
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    Tables &T = tables();
    std::lock_guard<std::mutex> lock(T.mutex);
    auto key = std::make_pair(dev, time);
    auto it = T.tables.find(key);
    if (it != T.tables.end()) return it->second;
    const int64_t rows = time + 1, count = rows * POISSON_BINS;
    std::vector<uint16_t> host(static_cast<size_t>(count));
    poisson_rows(host.data(), rows, POISSON_BINS);
    uint16_t *d = nullptr;
    SPK_HIP(hipMalloc(&d, size_t(count) * sizeof(uint16_t)));
    const hipError_t e = hipMemcpy(d, host.data(), size_t(count) * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        (void)hipFree(d);
        SPK_HIP(e);
    }
    T.tables.emplace(key, d);
    return d;
}

// mt + stream window (u32), times + values (+ rows) (u16)
size_t rate_lds(int64_t len, bool rows) { return size_t(4 * (MT_N + len + MT_N) + (rows ? 6 : 4) * len); }

template <typename O>
void launch_rate(const uint16_t *rows, const uint16_t *table, O *out, int64_t n, int64_t len, int32_t t,
                 uint64_t rnd, hipStream_t s)
{
    const bool in_lds = rate_lds(len, true) <= 131072;
    const size_t lds = rate_lds(len, in_lds);
    auto kernel = in_lds ? &rate_code_kernel<O, true> : &rate_code_kernel<O, false>;
    if (lds > 65536)
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)));
    kernel<<<dim3(unsigned(n)), RATE_THREADS, lds, s>>>(rows, table, out, len, t, rnd);
    SPK_LAUNCHED();
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_rate_code(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                      int64_t time, int32_t sort, void *stream)
{
    return guard([&] {
        SPK_ARG(time > 0, "Input time can't be less than 1.");
        SPK_ARG(time <= RATE_MAX_TIME, "RateCode: at most " << RATE_MAX_TIME << " time steps.");
        SPK_REQUIRE(in_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "RateCode: input data type " << type_name(in_type) << " is not supported.");
        SPK_REQUIRE(out_type == SPK_U8 || out_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "RateCode: output data type " << type_name(out_type) << " is not supported.");
        SPK_ARG(n >= 0 && len >= 1, "RateCode: input shape is invalid.");
        SPK_ARG(len <= RATE_MAX_LEN, "RateCode: at most " << RATE_MAX_LEN << " values per sample.");
        SPK_ARG(n < 65536, "RateCode: at most 65535 samples per call.");
        // one draw of the shared generator per call, as `auto rnd = Generator()` (code.cpp:195,214)
        const uint64_t rnd = generator_draw();
        if (n == 0) return;
        hipStream_t s = S(stream);
        const uint16_t *table = device_table(time);
        Scratch rows(n * len * int64_t(sizeof(uint16_t)), s);
        rank_times_launch((const float *)in, n, len, time, sort != 0, rows.as<uint16_t>(), s);
        if (out_type == SPK_U8)
            launch_rate<uint8_t>(rows.as<uint16_t>(), table, (uint8_t *)out, n, len, int32_t(time), rnd, s);
        else
            launch_rate<float>(rows.as<uint16_t>(), table, (float *)out, n, len, int32_t(time), rnd, s);
    });
}

int32_t spk_poisson_clear(void)
{
    return guard([&] {
        Tables &T = tables();
        std::lock_guard<std::mutex> lock(T.mutex);
        for (auto &kv : T.tables)
        {
            int cur = 0;
            SPK_HIP(hipGetDevice(&cur));
            SPK_HIP(hipSetDevice(kv.first.first));
            SPK_HIP(hipDeviceSynchronize());
            SPK_HIP(hipFree(kv.second));
            SPK_HIP(hipSetDevice(cur));
        }
        T.tables.clear();
    });
}
}
