// Rank-order coding (src/spyker/cpu/code.cpp:258-335).
//
// sort = true: every element of a sample gets the rank i of its value in
// descending order and the spike time U16(i * scale), scale = (T + 1e-5) /
// #{values above the sample minimum}.  Ranks come from a per-sample sort of
// 64-bit keys (descending-order value bits << 32 | index); keys are unique,
// so ties are ordered by ascending index.  The reference's std::sort leaves
// the order inside a tie block unspecified; the two agree whenever no tie
// block straddles a spike-time boundary.
//   1. chunks of CHUNK keys are sorted by a bitonic network (in registers,
//      wave shuffles and, for the largest strides, LDS);
//   2. sorted runs are merged pairwise until one run covers the sample; each
//      key finds its place by a binary search in the partner run;
//   3. each sorted position writes its element's time; 4. spikes are
//      (time <= t), written coalesced per step.
// sort = false: latency coding, time = U16(end - scale * (x - min)) with the
// subtraction fused as in the reference build (one fma).
#include "code.hpp"

namespace spk
{
namespace
{
constexpr int CHUNK = 2048;
constexpr int SORT_THREADS = 256;
constexpr int SORT_KPT = CHUNK / SORT_THREADS;  // keys a thread holds

__device__ __forceinline__ uint32_t desc_bits(float v)
{
    v = v + 0.0f;  // -0 -> +0: the reference's comparator treats them as equal
    uint32_t u = __float_as_uint(v);
    u ^= (u & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;  // ascending in v
    return ~u;                                             // descending in v
}

// float -> U16 as the reference's x86-64 build stores it: truncate to a
// 32-bit integer (NaN and out-of-range give 0x80000000) and keep 16 bits.
__device__ __forceinline__ uint16_t to_u16(float f)
{
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return 0;
    return uint16_t(uint32_t(int32_t(f)));
}

// 64-bit lane exchange within a wave (two 32-bit shuffles).
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int mask)
{
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), mask, 64));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), mask, 64));
    return (uint64_t(hi) << 32) | lo;
}

// One bitonic pass (size, STRIDE) over the thread's own SORT_KPT keys, at
// positions p0 + r: pairs (r, r + STRIDE) in the same thread.
template <int STRIDE>
__device__ __forceinline__ void pass_in_thread(uint64_t (&v)[SORT_KPT], int32_t p0, int32_t size)
{
#pragma unroll
    for (int r = 0; r < SORT_KPT; ++r)
        if ((r & STRIDE) == 0)
        {
            const bool up = ((p0 + r) & size) == 0;
            const uint64_t a = v[r], b = v[r + STRIDE];
            if ((a > b) == up) v[r] = b, v[r + STRIDE] = a;
        }
}

// The bitonic network of the original LDS form (pairs (i, i ^ stride),
// ascending where i & size is 0), with each thread holding the SORT_KPT
// consecutive positions p0 .. p0 + SORT_KPT - 1: strides below SORT_KPT stay in
// the thread, strides up to 64 * SORT_KPT / 2 exchange with a lane of the same
// wave, and only the larger ones (512 and 1024) pass through LDS, so a chunk
// takes 3 barrier pairs instead of one barrier per pass.  Keys are unique, so
// the sorted chunk does not depend on the network.
__global__ void __launch_bounds__(SORT_THREADS) chunk_sort(const float *__restrict__ in, uint64_t *__restrict__ keys,
                                                           int64_t len)
{
    __shared__ uint64_t k[CHUNK];
    const int64_t s = blockIdx.y, c0 = int64_t(blockIdx.x) * CHUNK;
    const float *src = in + s * len;
    const int32_t p0 = int32_t(threadIdx.x) * SORT_KPT;
    uint64_t v[SORT_KPT];
#pragma unroll
    for (int r = 0; r < SORT_KPT; ++r)
    {
        const int64_t j = c0 + p0 + r;
        v[r] = j < len ? (uint64_t(desc_bits(src[j])) << 32) | uint64_t(j) : ~0ull;
    }
    for (int32_t size = 2; size <= CHUNK; size <<= 1)
        for (int32_t stride = size >> 1; stride > 0; stride >>= 1)
        {
            if (stride < SORT_KPT)
            {
                if (stride == 4) pass_in_thread<4>(v, p0, size);
                else if (stride == 2) pass_in_thread<2>(v, p0, size);
                else pass_in_thread<1>(v, p0, size);
                continue;
            }
            uint64_t w[SORT_KPT];
            if (stride < 64 * SORT_KPT)
            {
#pragma unroll
                for (int r = 0; r < SORT_KPT; ++r) w[r] = shfl_xor64(v[r], stride / SORT_KPT);
            }
            else
            {
                __syncthreads();  // every read of the previous LDS pass is done
#pragma unroll
                for (int r = 0; r < SORT_KPT; ++r) k[p0 + r] = v[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < SORT_KPT; ++r) w[r] = k[(p0 + r) ^ stride];
            }
#pragma unroll
            for (int r = 0; r < SORT_KPT; ++r)
            {
                // the lower position of an ascending pair (or the upper of a
                // descending one) keeps the smaller key
                const int32_t p = p0 + r;
                const bool keep_min = ((p & stride) == 0) == ((p & size) == 0);
                const uint64_t a = v[r], b = w[r];
                v[r] = keep_min ? (a < b ? a : b) : (a < b ? b : a);
            }
        }
    uint64_t *dst = keys + s * len;
#pragma unroll
    for (int r = 0; r < SORT_KPT; ++r)
        if (c0 + p0 + r < len) dst[c0 + p0 + r] = v[r];
}

// Merge runs of length run into runs of length 2*run.
__global__ void merge_runs(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst, int64_t len, int64_t run)
{
    const int64_t s = blockIdx.y;
    const uint64_t *a = src + s * len;
    uint64_t *d = dst + s * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t r = i / run, p = i - r * run, q = r ^ 1;
        const int64_t qs = q * run, qe = min(qs + run, len);
        const uint64_t key = a[i];
        if (qs >= len)
        {
            d[i] = key;
            continue;
        }
        // number of keys in the partner run that are smaller than key
        int64_t lo = qs, hi = qe;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) >> 1;
            if (a[mid] < key)
                lo = mid + 1;
            else
                hi = mid;
        }
        d[min(r, q) * run + p + (lo - qs)] = key;
    }
}

// times[s][index] = U16(rank * scale)
__global__ void rank_times(const uint64_t *__restrict__ keys, uint16_t *__restrict__ times, int64_t len, int32_t t)
{
    const int64_t s = blockIdx.y;
    const uint64_t *k = keys + s * len;
    // elements strictly above the minimum = first sorted position whose value
    // equals the last (smallest) one
    const uint32_t last = uint32_t(k[len - 1] >> 32);
    int64_t lo = 0, hi = len;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if (uint32_t(k[mid] >> 32) < last)
            lo = mid + 1;
        else
            hi = mid;
    }
    const float scale = (float(t) + 1e-5f) / float(lo);
    uint16_t *ts = times + s * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x)
        ts[uint32_t(k[i])] = to_u16(float(int32_t(i)) * scale);
}

// merge_runs and rank_times for a sample of more than CHUNK and at most
// 2 * CHUNK keys (one merge pass, run = CHUNK): one block per sample holds its
// keys in LDS, each key finds its merged position by a binary search of the
// partner run there, the merged run replaces the two in LDS, and the spike
// times are written from it as rank_times writes them.
constexpr int MERGE_THREADS = 1024;
constexpr int MERGE_KPT = 2 * CHUNK / MERGE_THREADS;  // keys a thread places

__global__ void __launch_bounds__(MERGE_THREADS)
    merge_rank_lds(const uint64_t *__restrict__ src, uint16_t *__restrict__ times, int32_t len, int32_t t)
{
    __shared__ uint64_t k[2 * CHUNK];
    const int64_t s = blockIdx.x;
    const uint64_t *a = src + s * len;
    for (int32_t i = threadIdx.x; i < len; i += MERGE_THREADS) k[i] = a[i];
    __syncthreads();
    uint64_t key[MERGE_KPT];
    int32_t at[MERGE_KPT];
#pragma unroll
    for (int u = 0; u < MERGE_KPT; ++u)
    {
        const int32_t i = int32_t(threadIdx.x) + u * MERGE_THREADS;
        at[u] = -1;
        if (i >= len) continue;
        const int32_t r = i / CHUNK, p = i - r * CHUNK, q = r ^ 1;
        const int32_t qs = q * CHUNK, qe = min(qs + CHUNK, len);
        key[u] = k[i];
        // number of keys in the partner run that are smaller than key
        int32_t lo = qs, hi = qe;
        while (lo < hi)
        {
            const int32_t mid = (lo + hi) >> 1;
            if (k[mid] < key[u])
                lo = mid + 1;
            else
                hi = mid;
        }
        at[u] = min(r, q) * CHUNK + p + (lo - qs);
    }
    __syncthreads();  // every search is done before the merged run overwrites the two
#pragma unroll
    for (int u = 0; u < MERGE_KPT; ++u)
        if (at[u] >= 0) k[at[u]] = key[u];
    __syncthreads();
    // elements strictly above the minimum = first sorted position whose value
    // equals the last (smallest) one
    const uint32_t last = uint32_t(k[len - 1] >> 32);
    int32_t lo = 0, hi = len;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi) >> 1;
        if (uint32_t(k[mid] >> 32) < last)
            lo = mid + 1;
        else
            hi = mid;
    }
    const float scale = (float(t) + 1e-5f) / float(lo);
    uint16_t *ts = times + s * len;
    for (int32_t i = threadIdx.x; i < len; i += MERGE_THREADS) ts[uint32_t(k[i])] = to_u16(float(i) * scale);
}

__global__ void latency_times(const float *__restrict__ in, uint16_t *__restrict__ times, int64_t len, int32_t t)
{
    __shared__ float mm[32];
    const int64_t s = blockIdx.x;
    const float *x = in + s * len;
    float mn, mx;
    block_minmax(x, len, mn, mx, mm);
    const float end = float(t) + 1e-5f, scale = end / (mx - mn);
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x) times[s * len + i] = to_u16(__builtin_fmaf(-scale, x[i] - mn, end));
}

// out[s][step][j] = times[s][j] <= step; 4 bytes per thread per step.
template <typename O>
__global__ void spikes_from_times(const uint16_t *__restrict__ times, O *__restrict__ out, int64_t n, int64_t len,
                                  int32_t t)
{
    const int64_t total = n * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t s = i / len, j = i % len;
        const uint16_t tm = times[i];
        O *o = out + s * t * len + j;
        for (int32_t step = 0; step < t; ++step) o[int64_t(step) * len] = O(tm <= uint16_t(step));
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

namespace spk
{
void rank_times_launch(const float *in, int64_t n, int64_t len, int64_t time, bool sort, uint16_t *times,
                       hipStream_t s)
{
    if (sort)
    {
        Scratch k0(n * len * int64_t(sizeof(uint64_t)), s);
        const int64_t chunks = cdiv(len, CHUNK);
        chunk_sort<<<dim3(unsigned(chunks), unsigned(n)), SORT_THREADS, 0, s>>>(in, k0.as<uint64_t>(), len);
        SPK_LAUNCHED();
        if (len > CHUNK && len <= 2 * CHUNK)
        {
            // one merge pass and the times, the whole sample in LDS
            merge_rank_lds<<<dim3(unsigned(n)), MERGE_THREADS, 0, s>>>(k0.as<uint64_t>(), times, int32_t(len),
                                                                        int32_t(time));
            SPK_LAUNCHED();
            return;
        }
        // the second key buffer only for the global merge passes
        Scratch k1(len > CHUNK ? n * len * int64_t(sizeof(uint64_t)) : 0, s);
        uint64_t *src = k0.as<uint64_t>(), *dst = k1.as<uint64_t>();
        const unsigned blocks = unsigned(cdiv(len, 256) < 1024 ? cdiv(len, 256) : 1024);
        for (int64_t run = CHUNK; run < len; run <<= 1)
        {
            merge_runs<<<dim3(blocks, unsigned(n)), 256, 0, s>>>(src, dst, len, run);
            SPK_LAUNCHED();
            std::swap(src, dst);
        }
        rank_times<<<dim3(blocks, unsigned(n)), 256, 0, s>>>(src, times, len, int32_t(time));
        SPK_LAUNCHED();
    }
    else
    {
        latency_times<<<dim3(unsigned(n)), 1024, 0, s>>>(in, times, len, int32_t(time));
        SPK_LAUNCHED();
    }
}
}  // namespace spk

extern "C" {

int32_t spk_rank_code(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                      int64_t time, int32_t sort, void *stream)
{
    return guard([&] {
        SPK_ARG(time > 0, "Input time can't be less than 1.");
        SPK_ARG(time <= 65535, "Input time must fit the U16 spike times.");
        SPK_REQUIRE(in_type == SPK_F32, SPK_ERR_UNSUPPORTED, "RankCode: input data type " << type_name(in_type) << " is not supported.");
        SPK_REQUIRE(out_type == SPK_U8 || out_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "RankCode: output data type " << type_name(out_type) << " is not supported.");
        SPK_ARG(n >= 0 && len >= 1 && len < (1LL << 31), "RankCode: input shape is invalid.");
        SPK_ARG(n < 65536, "RankCode: at most 65535 samples per call.");
        if (n == 0) return;
        hipStream_t s = S(stream);
        Scratch times(n * len * int64_t(sizeof(uint16_t)), s);
        rank_times_launch((const float *)in, n, len, time, sort != 0, times.as<uint16_t>(), s);
        if (out_type == SPK_U8)
            spikes_from_times<uint8_t><<<flat_grid(n * len, 256), 256, 0, s>>>(times.as<uint16_t>(), (uint8_t *)out,
                                                                               n, len, int32_t(time));
        else
            spikes_from_times<float><<<flat_grid(n * len, 256), 256, 0, s>>>(times.as<uint16_t>(), (float *)out, n,
                                                                             len, int32_t(time));
        SPK_LAUNCHED();
    });
}
}
