// Sparse event path (src/spyker/sparse/*.cpp) on the GPU.
//
// A sparse tensor [n][t][c][h][w] is kept as one list per (sample, step) of
// the spikes that are new at that step, in CSR form: offsets[n*t + 1] and
// events packed z << 16 | y << 8 | x (the reference's U8 Spridx), each list
// in the reference's list order.  The producers:
//
//   - a list builder turns a first-spike-time map [n][c][h][w] into lists in
//     (z, y, x) scan order: per-chunk step histograms, their prefix sums,
//     and a stable scatter by one wave per chunk.  Rank code, dense->sparse
//     and conv all produce through it;
//   - pool and inhibit keep the first event of a cell / pixel (atomicMin of
//     the event's index, which orders events by step, then list order) and
//     filter the input lists stably (a scan of keep flags);
//   - convwta runs the reference's first-come loop, one wave per sample,
//     resolving a 64-event window at a time with ballots;
//   - conv keeps, per 16x16 output tile and chunk of 32 channels, the F32
//     potentials in registers; each step's events that reach the tile are
//     compacted in list order and added one by one, so every potential sees
//     the reference's sequence of F32 adds (sparse/conv.cpp:72-99);
//   - STDP builds a last-event-step map, one record of potentiate bits per
//     winner, and runs the per-channel update chains of the dense path with
//     the sparse path's F32 update rule.
#include "code.hpp"
#include "stdp.hpp"

namespace spk
{
namespace
{
constexpr int CHUNK = 1024;  // elements per builder chunk (one wave)
constexpr int CONV_TILE = 16;
constexpr int CONV_CC = 32;  // output channels per thread in the conv
constexpr int CONV_MAX_T = 1024;  // steps a conv block keeps the list offsets of

__device__ __forceinline__ int32_t pack(int32_t z, int32_t y, int32_t x) { return (z << 16) | (y << 8) | x; }
__device__ __forceinline__ int32_t ez(int32_t e) { return (e >> 16) & 0xFF; }
__device__ __forceinline__ int32_t ey(int32_t e) { return (e >> 8) & 0xFF; }
__device__ __forceinline__ int32_t ex(int32_t e) { return e & 0xFF; }

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

// -- scans ---------------------------------------------------------------------

constexpr int SCAN_THREADS = 256, SCAN_PER = 8, SCAN_SEG = SCAN_THREADS * SCAN_PER;

// Exclusive scan of each SCAN_SEG segment; segment totals to `sums`.
template <typename I>
__global__ void __launch_bounds__(SCAN_THREADS) scan_segments(const I *__restrict__ in, int64_t *__restrict__ out,
                                                              int64_t m, int64_t *__restrict__ sums)
{
    __shared__ int64_t part[SCAN_THREADS];
    const int64_t base = int64_t(blockIdx.x) * SCAN_SEG + int64_t(threadIdx.x) * SCAN_PER;
    int64_t v[SCAN_PER], run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) v[k] = base + k < m ? int64_t(in[base + k]) : 0, run += v[k];
    part[threadIdx.x] = run;
    __syncthreads();
    for (int d = 1; d < SCAN_THREADS; d <<= 1)
    {
        const int64_t o = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += o;
        __syncthreads();
    }
    int64_t acc = part[threadIdx.x] - run;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
    {
        if (base + k < m) out[base + k] = acc;
        acc += v[k];
    }
    if (threadIdx.x == SCAN_THREADS - 1) sums[blockIdx.x] = part[threadIdx.x];
}

__global__ void scan_add(int64_t *__restrict__ out, int64_t m, const int64_t *__restrict__ seg)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
        out[i] += seg[i / SCAN_SEG];
}

template <typename I>
__global__ void put_total(int64_t *__restrict__ out, int64_t m, const I *__restrict__ in)
{
    out[m] = m ? out[m - 1] + int64_t(in[m - 1]) : 0;
}

// out[0..m] = exclusive prefix sums of in[0..m), out[m] = the total.
template <typename I>
void exclusive_scan(const I *in, int64_t *out, int64_t m, hipStream_t s)
{
    if (m > 0)
    {
        const int64_t segs = cdiv(m, SCAN_SEG);
        Scratch sums(segs * 8, s), sums_scan((segs + 1) * 8, s);
        scan_segments<I><<<dim3(unsigned(segs)), SCAN_THREADS, 0, s>>>(in, out, m, sums.as<int64_t>());
        SPK_LAUNCHED();
        if (segs > 1)
        {
            exclusive_scan<int64_t>(sums.as<int64_t>(), sums_scan.as<int64_t>(), segs, s);
            scan_add<<<flat_grid(m, 256), 256, 0, s>>>(out, m, sums_scan.as<int64_t>());
            SPK_LAUNCHED();
        }
    }
    put_total<I><<<1, 1, 0, s>>>(out, m, in);
    SPK_LAUNCHED();
}

template <typename T>
__global__ void fill(T *__restrict__ p, int64_t m, T v)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <typename T>
void fill_launch(T *p, int64_t m, T v, hipStream_t s)
{
    if (m <= 0) return;
    fill<T><<<flat_grid(m, 256), 256, 0, s>>>(p, m, v);
    SPK_LAUNCHED();
}

int64_t read_total(const int64_t *p, hipStream_t s)
{
    int64_t v = 0;
    SPK_HIP(hipMemcpyAsync(&v, p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SPK_HIP(hipStreamSynchronize(s));
    return v;
}

// -- list builder -------------------------------------------------------------
//
// A first-spike-time map [n][len] (U16, >= t: no spike) becomes lists in
// scan order.  The elements of a sample are cut into chunks of CHUNK;
// cnt[(i*t + j)*chunks + c] counts chunk c's spikes at step j.  One exclusive
// scan of cnt in that order gives every (sample, step, chunk) its first
// position in the event array, and offsets[i*t + j] = pos[(i*t + j)*chunks].

__global__ void __launch_bounds__(256) chunk_hist(const uint16_t *__restrict__ times, int64_t len, int32_t t,
                                                  int64_t chunks, int32_t *__restrict__ cnt)
{
    extern __shared__ int32_t h[];  // [t]
    const int64_t i = blockIdx.y, c = blockIdx.x;
    for (int32_t k = threadIdx.x; k < t; k += 256) h[k] = 0;
    __syncthreads();
    const uint16_t *ts = times + i * len;
    const int64_t e1 = min(c * CHUNK + CHUNK, len);
    for (int64_t e = c * CHUNK + threadIdx.x; e < e1; e += 256)
    {
        const int32_t v = ts[e];
        if (v < t) atomicAdd(&h[v], 1);
    }
    __syncthreads();
    for (int32_t k = threadIdx.x; k < t; k += 256) cnt[(i * t + k) * chunks + c] = h[k];
}

__global__ void list_offsets(const int64_t *__restrict__ pos, int64_t chunks, int64_t lists, int64_t m,
                             int64_t *__restrict__ offsets)
{
    for (int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; l <= lists; l += int64_t(gridDim.x) * blockDim.x)
        offsets[l] = l < lists ? pos[l * chunks] : pos[m];
}

// Stable scatter of a chunk's spikes to their lists, one wave per chunk, 64
// elements a round: the lanes holding the same step find each other from
// one ballot per bit of the step (`bits` = bit width of t), so a round costs
// the same however many steps it holds.
__global__ void __launch_bounds__(64) chunk_scatter(const uint16_t *__restrict__ times, int64_t len, int32_t t,
                                                    int32_t bits, int64_t chunks, const int64_t *__restrict__ pos,
                                                    int32_t h, int32_t w, int32_t *__restrict__ events)
{
    extern __shared__ int64_t next[];  // [t] next position in each list
    const int64_t i = blockIdx.y, c = blockIdx.x;
    const int lane = threadIdx.x;
    for (int32_t k = lane; k < t; k += 64) next[k] = pos[(i * t + k) * chunks + c];
    __syncthreads();
    const uint16_t *ts = times + i * len;
    const int32_t hw = h * w;
    const int64_t e1 = min(c * CHUNK + CHUNK, len);
    for (int64_t e0 = c * CHUNK; e0 < e1; e0 += 64)
    {
        const int64_t e = e0 + lane;
        const int32_t v = e < e1 ? min(int32_t(ts[e]), t) : t;
        uint64_t same = ~0ull;
        for (int32_t b = 0; b < bits; ++b)
        {
            const uint64_t m = __ballot((v >> b) & 1);
            same &= ((v >> b) & 1) ? m : ~m;
        }
        const int32_t rank = __popcll(same & lanes_below(lane));
        const int64_t base = v < t ? next[v] : 0;
        __syncthreads();
        if (v < t)
        {
            const int32_t k = int32_t(e);
            events[base + rank] = pack(k / hw, (k / w) % h, k % w);
            if (rank == 0) next[v] = base + __popcll(same);
        }
        __syncthreads();
    }
}

// Lists of a [n][t][c][h][w] sparse tensor from its first-spike-time map
// [n][c*h*w] (U16, >= t: no spike), in (z, y, x) scan order.
void build_lists(const uint16_t *times, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, spk_sparse *out,
                 hipStream_t s)
{
    const int64_t len = c * h * w, chunks = cdiv(len, CHUNK), m = n * t * chunks;
    *out = spk_sparse{n, t, c, h, w, 0, nullptr, nullptr};
    void *p = nullptr;
    SPK_REQUIRE(spk_alloc(&p, (n * t + 1) * 8) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
    out->offsets = static_cast<int64_t *>(p);
    Scratch cnt(std::max<int64_t>(m, 1) * 4, s), pos((m + 1) * 8, s);
    if (m > 0)
    {
        chunk_hist<<<dim3(unsigned(chunks), unsigned(n)), 256, size_t(t) * 4, s>>>(times, len, int32_t(t), chunks,
                                                                                 cnt.as<int32_t>());
        SPK_LAUNCHED();
    }
    exclusive_scan<int32_t>(cnt.as<int32_t>(), pos.as<int64_t>(), m, s);
    list_offsets<<<flat_grid(n * t + 1, 256), 256, 0, s>>>(pos.as<int64_t>(), chunks, n * t, m, out->offsets);
    SPK_LAUNCHED();
    out->total = read_total(pos.as<int64_t>() + m, s);
    SPK_REQUIRE(spk_alloc(&p, std::max<int64_t>(out->total, 1) * 4) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
    out->events = static_cast<int32_t *>(p);
    if (out->total > 0)
    {
        const int32_t bits = 32 - __builtin_clz(uint32_t(t));
        chunk_scatter<<<dim3(unsigned(chunks), unsigned(n)), 64, size_t(t) * 8, s>>>(
            times, len, int32_t(t), bits, chunks, pos.as<int64_t>(), int32_t(h), int32_t(w), out->events);
        SPK_LAUNCHED();
    }
}

// -- dense <-> sparse, gather, pad ---------------------------------------------

template <typename T>
__global__ void first_over(const T *__restrict__ in, int64_t n, int32_t t, int64_t len, T threshold,
                           uint16_t *__restrict__ times)
{
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * len;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / len, k = idx % len;
        int32_t f = t;
        for (int32_t j = t - 1; j >= 0; --j) f = in[(i * t + j) * len + k] > threshold ? j : f;
        times[idx] = uint16_t(f);
    }
}

// Step of every event written at its element: block per list (i, j).
// Every producer emits an element at most once per sample (its first
// spike), so a plain store suffices and first and last step coincide.
__global__ void step_map(const int64_t *__restrict__ off, const int32_t *__restrict__ ev, int32_t t, int32_t h,
                         int32_t w, int64_t len, int32_t *__restrict__ map)
{
    const int64_t list = blockIdx.x, i = list / t;
    const int32_t j = int32_t(list % t);
    for (int64_t e = off[list] + threadIdx.x; e < off[list + 1]; e += blockDim.x)
    {
        const int32_t v = ev[e];
        map[i * len + (int64_t(ez(v)) * h + ey(v)) * w + ex(v)] = j;
    }
}

template <typename O>
__global__ void map_out(const int32_t *__restrict__ map, O *__restrict__ out, int64_t m)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = O(map[i]);
}

// Cumulative spikes from the first-step map: a thread reads the steps of 4
// consecutive elements once and writes their 4 bytes at every step (one
// dword store when `vec`: len % 4 == 0 and out 4-byte aligned).
__global__ void dense_from_map(const int32_t *__restrict__ map, uint8_t *__restrict__ out, int64_t n, int32_t t,
                               int64_t len, int32_t vec)
{
    const int64_t groups = (len + 3) / 4;
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * groups;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / groups, k0 = (idx % groups) * 4;
        const int32_t *m = map + i * len + k0;
        int32_t f[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) f[q] = k0 + q < len ? m[q] : t;
        uint8_t *o = out + i * t * len + k0;
        for (int32_t j = 0; j < t; ++j, o += len)
        {
            if (vec)
                *reinterpret_cast<uint32_t *>(o) = uint32_t(f[0] <= j) | uint32_t(f[1] <= j) << 8 |
                                                   uint32_t(f[2] <= j) << 16 | uint32_t(f[3] <= j) << 24;
            else
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (k0 + q < len) o[q] = f[q] <= j;
        }
    }
}

// fcwta on the spike lists: the winners of fcwta(dense(), radius, count)
// over the flat index (z*h + y)*w + x.  Candidates come in (step, flat
// index) order -- the dense branch's order of T - first step, descending,
// then index -- and one is a winner unless an earlier winner lies within
// `radius` of it.  One wave per sample; a list already in ascending flat
// order (what every scan-order producer emits) is walked 64 events at a
// time, any other list by repeated minimum selection.
__global__ void __launch_bounds__(64) sparse_fcwta(const int64_t *__restrict__ off, const int32_t *__restrict__ ev,
                                                   int32_t t, int32_t h, int32_t w, int32_t radius, int32_t count,
                                                   spk_winner *__restrict__ win, int64_t *__restrict__ counts)
{
    extern __shared__ int32_t wflat[];  // [count]
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x;
    auto flat = [&](int32_t v) { return (ez(v) * h + ey(v)) * w + ex(v); };
    auto excluded = [&](int32_t f, int32_t found) {
        bool x = false;
        for (int32_t k = 0; k < found; ++k) x |= abs(f - wflat[k]) <= radius;
        return x;
    };
    auto add = [&](int32_t v, int32_t j, int32_t found) {
        if (lane == 0)
        {
            win[i * count + found] = spk_winner{0, j, ez(v), ey(v), ex(v)};
            wflat[found] = flat(v);
        }
        __syncthreads();
    };
    int32_t found = 0;
    for (int32_t j = 0; j < t && found < count; ++j)
    {
        const int64_t e0 = off[i * t + j], e1 = off[i * t + j + 1];
        bool sorted = true;
        for (int64_t b = e0 + 1; b < e1 && sorted; b += 64)
        {
            const int64_t e = b + lane;
            sorted = !__ballot(e < e1 && flat(ev[e - 1]) >= flat(ev[e]));
        }
        if (sorted)
        {
            for (int64_t b = e0; b < e1 && found < count; b += 64)
            {
                const int64_t e = b + lane;
                const int32_t v = e < e1 ? ev[e] : 0, f = flat(v);
                bool ok = e < e1 && !excluded(f, found);
                uint64_t pass = __ballot(ok);
                while (pass && found < count)
                {
                    const int l = __builtin_ctzll(pass);
                    const int32_t vl = __shfl(v, l, 64);
                    add(vl, j, found);
                    ++found;
                    if (abs(f - flat(vl)) <= radius) ok = false;
                    pass = __ballot(ok) & (l == 63 ? 0ull : ~0ull << (l + 1));
                }
            }
        }
        else
            while (found < count)
            {
                int32_t best = INT32_MAX, bv = 0;
                for (int64_t b = e0; b < e1; b += 64)
                {
                    const int64_t e = b + lane;
                    if (e < e1)
                    {
                        const int32_t v = ev[e], f = flat(v);
                        if (f < best && !excluded(f, found)) best = f, bv = v;
                    }
                }
                for (int d = 32; d >= 1; d >>= 1)
                {
                    const int32_t ob = __shfl_xor(best, d, 64), ov = __shfl_xor(bv, d, 64);
                    if (ob < best) best = ob, bv = ov;
                }
                if (best == INT32_MAX) break;
                add(bv, j, found);
                ++found;
            }
    }
    if (lane == 0) counts[i] = found;
}

__global__ void shift_events(const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t m, int32_t dy,
                             int32_t dx)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = in[i] + (dy << 8) + dx;
}

// -- pool, inhibit: first event per cell, then a stable filter ----------------------

// mode 0 (pool): cell = (z, (pt + y) / kh, (pl + x) / kw), dropped past the
// output; mode 1 (inhibit): cell = the pixel.  first[i][cell] = the lowest
// index (within the sample) of an event reaching it.
__global__ void first_event(const int64_t *__restrict__ off, const int32_t *__restrict__ ev, int32_t t, int32_t mode,
                            int32_t kh, int32_t kw, int32_t pt, int32_t pl, int32_t ch, int32_t cw, int64_t cells,
                            int32_t *__restrict__ first)
{
    const int64_t list = blockIdx.x, i = list / t, s0 = off[i * t];
    for (int64_t e = off[list] + threadIdx.x; e < off[list + 1]; e += blockDim.x)
    {
        const int32_t v = ev[e];
        int64_t cell;
        if (mode == 0)
        {
            const int32_t y = (pt + ey(v)) / kh, x = (pl + ex(v)) / kw;
            if (y >= ch || x >= cw) continue;
            cell = (int64_t(ez(v)) * ch + y) * cw + x;
        }
        else
            cell = int64_t(ey(v)) * cw + ex(v);
        atomicMin(first + i * cells + cell, int32_t(e - s0));
    }
}

__global__ void keep_first(const int64_t *__restrict__ off, const int32_t *__restrict__ ev, int32_t t, int32_t mode,
                           int32_t kh, int32_t kw, int32_t pt, int32_t pl, int32_t ch, int32_t cw, int64_t cells,
                           const int32_t *__restrict__ first, int32_t *__restrict__ flag, int32_t *__restrict__ value)
{
    const int64_t list = blockIdx.x, i = list / t, s0 = off[i * t];
    for (int64_t e = off[list] + threadIdx.x; e < off[list + 1]; e += blockDim.x)
    {
        const int32_t v = ev[e];
        int32_t keep = 0, out = v;
        if (mode == 0)
        {
            const int32_t y = (pt + ey(v)) / kh, x = (pl + ex(v)) / kw;
            if (y < ch && x < cw)
            {
                keep = first[i * cells + (int64_t(ez(v)) * ch + y) * cw + x] == int32_t(e - s0);
                out = pack(ez(v), y, x);
            }
        }
        else
        {
            const int32_t f = first[i * cells + int64_t(ey(v)) * cw + ex(v)];
            keep = ez(ev[s0 + f]) == ez(v);
        }
        flag[e] = keep;
        value[e] = out;
    }
}

__global__ void filter_scatter(const int32_t *__restrict__ flag, const int32_t *__restrict__ value,
                               const int64_t *__restrict__ pos, int64_t m, int32_t *__restrict__ out)
{
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += int64_t(gridDim.x) * blockDim.x)
        if (flag[e]) out[pos[e]] = value[e];
}

__global__ void filter_offsets(const int64_t *__restrict__ off, const int64_t *__restrict__ pos, int64_t lists,
                               int64_t *__restrict__ out)
{
    for (int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; l <= lists; l += int64_t(gridDim.x) * blockDim.x)
        out[l] = pos[off[l]];
}

// Keeps the flagged events (with their new values) in order.
void filter_lists(const spk_sparse *in, const int32_t *flag, const int32_t *value, spk_sparse *out, hipStream_t s)
{
    const int64_t lists = in->n * in->t, m = in->total;
    Scratch pos((m + 1) * 8, s);
    exclusive_scan<int32_t>(flag, pos.as<int64_t>(), m, s);
    void *p = nullptr;
    SPK_REQUIRE(spk_alloc(&p, (lists + 1) * 8) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
    out->offsets = static_cast<int64_t *>(p);
    filter_offsets<<<flat_grid(lists + 1, 256), 256, 0, s>>>(in->offsets, pos.as<int64_t>(), lists, out->offsets);
    SPK_LAUNCHED();
    out->total = read_total(pos.as<int64_t>() + m, s);
    SPK_REQUIRE(spk_alloc(&p, std::max<int64_t>(out->total, 1) * 4) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
    out->events = static_cast<int32_t *>(p);
    if (m > 0)
    {
        filter_scatter<<<flat_grid(m, 256), 256, 0, s>>>(flag, value, pos.as<int64_t>(), m, out->events);
        SPK_LAUNCHED();
    }
}

// -- convwta: the reference's first-come loop, one wave per sample ------------------

__global__ void __launch_bounds__(64) sparse_wta(const int64_t *__restrict__ off, const int32_t *__restrict__ ev,
                                                 int32_t t, int32_t h, int32_t w, int32_t ry, int32_t rx,
                                                 int32_t count, spk_winner *__restrict__ win,
                                                 int64_t *__restrict__ counts)
{
    extern __shared__ int32_t box[];  // [count][5]: z, ys, ye, xs, xe
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x;
    int32_t found = 0;
    for (int32_t j = 0; j < t && found < count; ++j)
        for (int64_t e0 = off[i * t + j]; e0 < off[i * t + j + 1] && found < count; e0 += 64)
        {
            const int64_t e = e0 + lane;
            const bool valid = e < off[i * t + j + 1];
            const int32_t v = valid ? ev[e] : 0, z = ez(v), y = ey(v), x = ex(v);
            bool ok = valid;
            for (int32_t k = 0; k < found; ++k)
            {
                const int32_t *b = box + 5 * k;
                if (z == b[0] || (b[1] <= y && y <= b[2] && b[3] <= x && x <= b[4])) ok = false;
            }
            uint64_t pass = __ballot(ok);
            while (pass && found < count)
            {
                const int l = __builtin_ctzll(pass);
                const int32_t zl = __shfl(z, l, 64), yl = __shfl(y, l, 64), xl = __shfl(x, l, 64);
                const int32_t ys = max(0, yl - ry), ye = min(h - 1, yl + ry);
                const int32_t xs = max(0, xl - rx), xe = min(w - 1, xl + rx);
                if (lane == 0)
                {
                    win[i * count + found] = spk_winner{0, j, zl, yl, xl};
                    int32_t *b = box + 5 * found;
                    b[0] = zl, b[1] = ys, b[2] = ye, b[3] = xs, b[4] = xe;
                }
                ++found;
                __syncthreads();
                if (z == zl || (ys <= y && y <= ye && xs <= x && x <= xe)) ok = false;
                pass = __ballot(ok) & ~(l == 63 ? ~0ull : (~0ull >> (63 - l)));
            }
        }
    if (lane == 0) counts[i] = found;
}

// -- conv ------------------------------------------------------------------------

struct SparseConvGeom
{
    int32_t t, cin, h, w, cout, kh, kw, pt, pl, ho, wo, tiles_x;
    float threshold;
};

// One 16x16 output tile of one sample and CONV_CC output channels per block,
// a thread per pixel with its CONV_CC potentials in registers; the weights
// sit in LDS as [z][ky][kx][CONV_CC] so a tap's channels are 8 float4 reads.
// The sample's events stream through in rounds of 256 (the next round's
// loads in flight while this one is applied); the events of a round that
// reach the tile are compacted in list order with their step, and every
// potential gets its adds in the reference's order.  The threshold check
// runs where the tile's hits change step, after the last hit, and after the
// sample's first step with events: a potential that no event reached since
// the last check cannot newly exceed the threshold, except the initial 0
// when the threshold is negative, which that first check covers.  Writes the
// first spike step of every neuron (t where none).
__global__ void __launch_bounds__(256) sparse_conv_tile(const int64_t *__restrict__ off,
                                                        const int32_t *__restrict__ ev,
                                                        const float *__restrict__ kernel, SparseConvGeom g,
                                                        uint16_t *__restrict__ times)
{
    extern __shared__ float4 wl4[];  // [cin][kh][kw][CONV_CC / 4]
    __shared__ int32_t elist[256], estep[256];
    __shared__ int32_t wave_n[4];
    __shared__ int64_t loff[CONV_MAX_T + 1];
    float *wl = reinterpret_cast<float *>(wl4);
    const int64_t i = blockIdx.y;
    const int32_t c0 = blockIdx.z * CONV_CC;
    const int32_t ty0 = (blockIdx.x / g.tiles_x) * CONV_TILE, tx0 = (blockIdx.x % g.tiles_x) * CONV_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t oy = ty0 + tid / CONV_TILE, ox = tx0 + tid % CONV_TILE;
    const bool pix = oy < g.ho && ox < g.wo;
    const int32_t taps = g.cin * g.kh * g.kw;
    for (int32_t q = tid; q < CONV_CC * taps; q += 256)
    {
        const int32_t c = c0 + q % CONV_CC;
        wl[q] = c < g.cout ? kernel[int64_t(c) * taps + q / CONV_CC] : 0.0f;
    }
    for (int32_t k = tid; k <= g.t; k += 256) loff[k] = off[i * g.t + k];
    float pot[CONV_CC];
    int32_t fire[CONV_CC];
#pragma unroll
    for (int c = 0; c < CONV_CC; ++c) pot[c] = 0.0f, fire[c] = g.t;
    __syncthreads();
    auto check = [&](int32_t j) {
#pragma unroll
        for (int c = 0; c < CONV_CC; ++c)
            if (pot[c] > g.threshold) fire[c] = j, pot[c] = -FLT_MAX;
    };
    const int64_t s0 = loff[0], s1 = loff[g.t];
    // step of the last applied hits, starting at the sample's first step with
    // events so that step is checked after its hits even if none reach here
    int32_t cur = -1;
    if (s0 < s1)
    {
        cur = 0;
        while (loff[cur + 1] == s0) ++cur;
    }
    int32_t vn = s0 + tid < s1 ? ev[s0 + tid] : -1;
    for (int64_t r0 = s0; r0 < s1; r0 += 256)
    {
        const int64_t e = r0 + tid;
        const int32_t v = vn;
        vn = e + 256 < s1 ? ev[e + 256] : -1;  // next round, in flight meanwhile
        const int32_t dy = v >= 0 ? ey(v) + g.pt - ty0 : -1, dx = v >= 0 ? ex(v) + g.pl - tx0 : -1;
        const bool hit = v >= 0 && dy >= 0 && dy < CONV_TILE + g.kh - 1 && dx >= 0 && dx < CONV_TILE + g.kw - 1;
        const uint64_t b = __ballot(hit);
        if (lane == 0) wave_n[wave] = __popcll(b);
        __syncthreads();
        int32_t at = __popcll(b & lanes_below(lane)), total = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) at += k < wave ? wave_n[k] : 0, total += wave_n[k];
        if (hit)
        {
            // step of event e: the last k with loff[k] <= e
            int32_t lo = 0, hi = g.t - 1;
            while (lo < hi)
            {
                const int32_t mid = (lo + hi + 1) >> 1;
                if (loff[mid] <= e)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            elist[at] = v, estep[at] = lo;
        }
        __syncthreads();
        for (int32_t q = 0; q < total; ++q)
        {
            const int32_t u = elist[q], j = estep[q];
            if (j != cur) check(cur), cur = j;
            const int32_t ky = ey(u) + g.pt - oy, kx = ex(u) + g.pl - ox;
            if (pix && ky >= 0 && ky < g.kh && kx >= 0 && kx < g.kw)
            {
                const float4 *wk = wl4 + ((ez(u) * g.kh + ky) * g.kw + kx) * (CONV_CC / 4);
#pragma unroll
                for (int c = 0; c < CONV_CC / 4; ++c)
                {
                    const float4 w4 = wk[c];
                    pot[4 * c] += w4.x, pot[4 * c + 1] += w4.y, pot[4 * c + 2] += w4.z, pot[4 * c + 3] += w4.w;
                }
            }
        }
        __syncthreads();
    }
    if (cur >= 0) check(cur);
    if (pix)
#pragma unroll
        for (int c = 0; c < CONV_CC; ++c)
            if (c0 + c < g.cout) times[((i * g.cout + c0 + c) * g.ho + oy) * g.wo + ox] = uint16_t(fire[c]);
}

// -- STDP records ---------------------------------------------------------------

// One wave per winner slot: slot key and the potentiate bits of its kernel
// row from the last-event map (-1: none).
__global__ void __launch_bounds__(256) sparse_records(const int32_t *__restrict__ last, StdpGeom g,
                                                      const spk_winner *__restrict__ win,
                                                      const int64_t *__restrict__ counts, uint32_t *__restrict__ bits,
                                                      int2 *__restrict__ slot, int32_t *__restrict__ flag)
{
    const int64_t e = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (e >= g.n * g.maxcount) return;
    const int lane = threadIdx.x & 63;
    const int64_t s = e / g.maxcount, k = e % g.maxcount;
    const spk_winner w = k < counts[s] ? win[e] : spk_winner{-1, 0, 0, 0, 0};
    int32_t bad = 0;
    if (w.c >= g.ncfg)
        bad = 1;
    else if (w.c >= 0 && (w.z < 0 || w.z >= g.cout || w.y < 0 || w.y >= g.ho || w.x < 0 || w.x >= g.wo))
        bad = 2;
    if (lane == 0)
    {
        slot[e] = make_int2(w.c >= 0 && !bad ? int32_t(w.z) : -1, int32_t(w.c));
        if (bad) atomicMax(flag, bad);
    }
    if (w.c < 0 || bad) return;
    const int32_t taps = g.kh * g.kw, ksz = g.cin * taps, words = (ksz + 31) / 32;
    const int32_t *ls = last + s * int64_t(g.cin) * g.h * g.w;
    uint32_t *rec = bits + e * words;
    for (int32_t base = 0; base < ksz; base += 64)
    {
        const int32_t q = base + lane;
        bool bit = false;
        if (q < ksz)
        {
            const int32_t z = q / taps, iy = int32_t(w.y) + (q / g.kw) % g.kh - g.pt, ix = int32_t(w.x) + q % g.kw - g.pl;
            int32_t step = g.t;
            if (iy >= 0 && iy < g.h && ix >= 0 && ix < g.w)
            {
                const int32_t l = ls[(int64_t(z) * g.h + iy) * g.w + ix];
                step = l >= 0 ? l : g.t;
            }
            bit = step <= w.t;
        }
        const uint64_t b = __ballot(bit);
        if (lane == 0 && base / 32 < words) rec[base / 32] = uint32_t(b);
        if (lane == 32 && base / 32 + 1 < words) rec[base / 32 + 1] = uint32_t(b >> 32);
    }
}

__global__ void clear_flag(int32_t *flag) { *flag = 0; }

void check_sparse(const spk_sparse *in)
{
    SPK_ARG(in && in->offsets && in->events, "Sparse tensor is not initialized.");
    SPK_ARG(in->n >= 0 && in->t >= 1 && in->c >= 1 && in->h >= 1 && in->w >= 1, "Sparse tensor shape is invalid.");
    SPK_ARG(in->c <= 255 && in->h <= 255 && in->w <= 255, "Sparse tensors hold at most 255 channels, rows and columns.");
    SPK_ARG(in->t <= 4096 && in->n < 65536, "Sparse tensors hold at most 4096 steps and 65535 samples.");
    SPK_ARG(in->total < INT32_MAX, "Sparse tensors hold fewer than 2^31 events.");
}

// first (mode 0) or last (mode 1) event step per element, [n][c][h][w] int32;
// none: t (first) or -1 (last)
void step_map_launch(const spk_sparse *in, int32_t *map, int32_t mode, hipStream_t s)
{
    const int64_t len = in->c * in->h * in->w, lists = in->n * in->t;
    fill_launch<int32_t>(map, in->n * len, mode == 0 ? int32_t(in->t) : -1, s);
    if (lists > 0 && in->total > 0)
    {
        step_map<<<dim3(unsigned(lists)), 256, 0, s>>>(in->offsets, in->events, int32_t(in->t), int32_t(in->h),
                                                       int32_t(in->w), len, map);
        SPK_LAUNCHED();
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_sparse_release(spk_sparse *s)
{
    return guard([&] {
        SPK_ARG(s, "Sparse tensor is null.");
        if (s->offsets) SPK_REQUIRE(spk_dealloc(s->offsets) == SPK_OK, SPK_ERR_ARGUMENT, spk_last_error());
        if (s->events) SPK_REQUIRE(spk_dealloc(s->events) == SPK_OK, SPK_ERR_ARGUMENT, spk_last_error());
        s->offsets = nullptr, s->events = nullptr, s->total = 0;
    });
}

int32_t spk_sparse_from_dense(const void *in, int32_t type, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                              double threshold, spk_sparse *out, void *stream)
{
    return guard([&] {
        SPK_ARG(out, "Sparse output is null.");
        SPK_ARG(n >= 0 && t >= 1 && c >= 1 && h >= 1 && w >= 1, "Sparse tensor shape is invalid.");
        SPK_ARG(c <= 255 && h <= 255 && w <= 255, "Sparse tensors hold at most 255 channels, rows and columns.");
        SPK_ARG(t <= 4096 && n < 65536, "Sparse tensors hold at most 4096 steps and 65535 samples.");
        SPK_REQUIRE(type == SPK_U8 || type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "Sparse: input data type " << type_name(type) << " is not supported.");
        hipStream_t s = S(stream);
        const int64_t len = c * h * w;
        Scratch times(std::max<int64_t>(n * len, 1) * 2, s);
        if (n > 0)
        {
            if (type == SPK_U8)
                first_over<uint8_t><<<flat_grid(n * len, 256), 256, 0, s>>>(
                    (const uint8_t *)in, n, int32_t(t), len, uint8_t(threshold), times.as<uint16_t>());
            else
                first_over<float><<<flat_grid(n * len, 256), 256, 0, s>>>((const float *)in, n, int32_t(t), len,
                                                                          float(threshold), times.as<uint16_t>());
            SPK_LAUNCHED();
        }
        build_lists(times.as<uint16_t>(), n, t, c, h, w, out, s);
    });
}

int32_t spk_sparse_code(const void *in, int32_t type, int64_t n, int64_t c, int64_t h, int64_t w, int64_t time,
                        int32_t sort, spk_sparse *out, void *stream)
{
    return guard([&] {
        SPK_ARG(out, "Sparse output is null.");
        SPK_ARG(time > 0, "Input time can't be less than 1.");
        SPK_ARG(time <= 4096, "Sparse tensors hold at most 4096 steps and 65535 samples.");
        SPK_ARG(n >= 0 && n < 65536 && c >= 1 && h >= 1 && w >= 1, "Sparse code: input shape is invalid.");
        SPK_ARG(c <= 255 && h <= 255 && w <= 255, "Sparse tensors hold at most 255 channels, rows and columns.");
        SPK_REQUIRE(type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "Sparse code: input data type " << type_name(type) << " is not supported.");
        hipStream_t s = S(stream);
        const int64_t len = c * h * w;
        Scratch times(std::max<int64_t>(n * len, 1) * 2, s);
        if (n > 0) rank_times_launch((const float *)in, n, len, time, sort != 0, times.as<uint16_t>(), s);
        build_lists(times.as<uint16_t>(), n, time, c, h, w, out, s);
    });
}

int32_t spk_sparse_to_dense(const spk_sparse *in, void *out, void *stream)
{
    return guard([&] {
        check_sparse(in);
        hipStream_t s = S(stream);
        const int64_t len = in->c * in->h * in->w;
        Scratch map(std::max<int64_t>(in->n * len, 1) * 4, s);
        step_map_launch(in, map.as<int32_t>(), 0, s);
        if (in->n * in->t * len > 0)
        {
            const int32_t vec = len % 4 == 0 && reinterpret_cast<uintptr_t>(out) % 4 == 0;
            dense_from_map<<<flat_grid(in->n * cdiv(len, 4), 256), 256, 0, s>>>(map.as<int32_t>(), (uint8_t *)out,
                                                                                in->n, int32_t(in->t), len, vec);
            SPK_LAUNCHED();
        }
    });
}

int32_t spk_sparse_gather(const spk_sparse *in, void *out, int32_t type, void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_REQUIRE(type == SPK_U8 || type == SPK_U16 || type == SPK_I32 || type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "Sparse gather: output data type " << type_name(type) << " is not supported.");
        SPK_ARG(type != SPK_U8 || in->t <= 255, "Sparse gather: U8 output holds at most 255 steps.");
        hipStream_t s = S(stream);
        const int64_t m = in->n * in->c * in->h * in->w;
        Scratch map(std::max<int64_t>(m, 1) * 4, s);
        step_map_launch(in, map.as<int32_t>(), 0, s);
        if (m == 0) return;
        const dim3 grid = flat_grid(m, 256);
        if (type == SPK_U8) map_out<uint8_t><<<grid, 256, 0, s>>>(map.as<int32_t>(), (uint8_t *)out, m);
        if (type == SPK_U16) map_out<uint16_t><<<grid, 256, 0, s>>>(map.as<int32_t>(), (uint16_t *)out, m);
        if (type == SPK_I32) map_out<int32_t><<<grid, 256, 0, s>>>(map.as<int32_t>(), (int32_t *)out, m);
        if (type == SPK_F32) map_out<float><<<grid, 256, 0, s>>>(map.as<int32_t>(), (float *)out, m);
        SPK_LAUNCHED();
    });
}

int32_t spk_sparse_pad(const spk_sparse *in, int64_t pt, int64_t pl, int64_t pb, int64_t pr, spk_sparse *out,
                       void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(out, "Sparse output is null.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        SPK_ARG(in->h + pt + pb <= 255 && in->w + pl + pr <= 255,
                "Sparse tensors hold at most 255 channels, rows and columns.");
        hipStream_t s = S(stream);
        const int64_t lists = in->n * in->t;
        *out = spk_sparse{in->n, in->t, in->c, in->h + pt + pb, in->w + pl + pr, in->total, nullptr, nullptr};
        void *p = nullptr;
        SPK_REQUIRE(spk_alloc(&p, (lists + 1) * 8) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
        out->offsets = static_cast<int64_t *>(p);
        SPK_REQUIRE(spk_alloc(&p, std::max<int64_t>(in->total, 1) * 4) == SPK_OK, SPK_ERR_RUNTIME, spk_last_error());
        out->events = static_cast<int32_t *>(p);
        SPK_HIP(hipMemcpyAsync(out->offsets, in->offsets, size_t(lists + 1) * 8, hipMemcpyDeviceToDevice, s));
        if (in->total > 0)
        {
            shift_events<<<flat_grid(in->total, 256), 256, 0, s>>>(in->events, out->events, in->total, int32_t(pt),
                                                                   int32_t(pl));
            SPK_LAUNCHED();
        }
    });
}

int32_t spk_sparse_pool(const spk_sparse *in, int64_t kh, int64_t kw, int64_t sh, int64_t sw, int64_t pt, int64_t pl,
                        int64_t pb, int64_t pr, spk_sparse *out, void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(out, "Sparse output is null.");
        SPK_ARG(kh >= 1 && kw >= 1, "Kernel items are incorrect.");
        SPK_ARG(sh == kh && sw == kw, "Stride must be equal to kernel in sparse pool.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        const int64_t ho = (in->h + pt + pb - kh) / kh + 1, wo = (in->w + pl + pr - kw) / kw + 1;
        SPK_ARG(in->h + pt + pb >= kh && in->w + pl + pr >= kw, "Input and kernel shapes are not compatible.");
        hipStream_t s = S(stream);
        const int64_t cells = in->c * ho * wo, lists = in->n * in->t;
        *out = spk_sparse{in->n, in->t, in->c, ho, wo, 0, nullptr, nullptr};
        Scratch first(std::max<int64_t>(in->n * cells, 1) * 4, s), flag(std::max<int64_t>(in->total, 1) * 4, s),
            value(std::max<int64_t>(in->total, 1) * 4, s);
        fill_launch<int32_t>(first.as<int32_t>(), in->n * cells, INT32_MAX, s);
        if (in->total > 0)
        {
            first_event<<<dim3(unsigned(lists)), 256, 0, s>>>(in->offsets, in->events, int32_t(in->t), 0, int32_t(kh),
                                                              int32_t(kw), int32_t(pt), int32_t(pl), int32_t(ho),
                                                              int32_t(wo), cells, first.as<int32_t>());
            SPK_LAUNCHED();
            keep_first<<<dim3(unsigned(lists)), 256, 0, s>>>(in->offsets, in->events, int32_t(in->t), 0, int32_t(kh),
                                                             int32_t(kw), int32_t(pt), int32_t(pl), int32_t(ho),
                                                             int32_t(wo), cells, first.as<int32_t>(),
                                                             flag.as<int32_t>(), value.as<int32_t>());
            SPK_LAUNCHED();
        }
        filter_lists(in, flag.as<int32_t>(), value.as<int32_t>(), out, s);
    });
}

int32_t spk_sparse_inhibit(const spk_sparse *in, spk_sparse *out, void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(out, "Sparse output is null.");
        hipStream_t s = S(stream);
        const int64_t cells = in->h * in->w, lists = in->n * in->t;
        *out = spk_sparse{in->n, in->t, in->c, in->h, in->w, 0, nullptr, nullptr};
        Scratch first(std::max<int64_t>(in->n * cells, 1) * 4, s), flag(std::max<int64_t>(in->total, 1) * 4, s),
            value(std::max<int64_t>(in->total, 1) * 4, s);
        fill_launch<int32_t>(first.as<int32_t>(), in->n * cells, INT32_MAX, s);
        if (in->total > 0)
        {
            first_event<<<dim3(unsigned(lists)), 256, 0, s>>>(in->offsets, in->events, int32_t(in->t), 1, 1, 1, 0, 0,
                                                              int32_t(in->h), int32_t(in->w), cells,
                                                              first.as<int32_t>());
            SPK_LAUNCHED();
            keep_first<<<dim3(unsigned(lists)), 256, 0, s>>>(in->offsets, in->events, int32_t(in->t), 1, 1, 1, 0, 0,
                                                             int32_t(in->h), int32_t(in->w), cells,
                                                             first.as<int32_t>(), flag.as<int32_t>(),
                                                             value.as<int32_t>());
            SPK_LAUNCHED();
        }
        filter_lists(in, flag.as<int32_t>(), value.as<int32_t>(), out, s);
    });
}

int32_t spk_sparse_convwta(const spk_sparse *in, int64_t ry, int64_t rx, int64_t count, spk_winner *winners,
                           int64_t *counts, void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(ry >= 0 && rx >= 0, "Radius items are incorrect.");
        SPK_ARG(count > 0, "Winners count can't be less than 1.");
        SPK_ARG(count <= 3072, "Sparse convwta: at most 3072 winners per sample.");
        SPK_ARG(winners && counts, "Sparse convwta: winner buffers are required.");
        if (in->n == 0) return;
        hipStream_t s = S(stream);
        sparse_wta<<<dim3(unsigned(in->n)), 64, size_t(count) * 20, s>>>(
            in->offsets, in->events, int32_t(in->t), int32_t(in->h), int32_t(in->w), int32_t(ry), int32_t(rx),
            int32_t(count), winners, counts);
        SPK_LAUNCHED();
    });
}

int32_t spk_sparse_fcwta(const spk_sparse *in, int64_t radius, int64_t count, spk_winner *winners, int64_t *counts,
                         void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(radius >= 0, "Radius can't be less than 0.");
        SPK_ARG(count > 0, "Winners count can't be less than 1.");
        SPK_ARG(count <= 8192, "Sparse fcwta: at most 8192 winners per sample.");
        SPK_ARG(winners && counts, "Sparse fcwta: winner buffers are required.");
        if (in->n == 0) return;
        hipStream_t s = S(stream);
        sparse_fcwta<<<dim3(unsigned(in->n)), 64, size_t(count) * 4, s>>>(
            in->offsets, in->events, int32_t(in->t), int32_t(in->h), int32_t(in->w),
            int32_t(std::min<int64_t>(radius, INT32_MAX)), int32_t(count), winners, counts);
        SPK_LAUNCHED();
    });
}

int32_t spk_sparse_conv(const spk_sparse *in, const void *kernel, int32_t kernel_type, int64_t cout, int64_t kh,
                        int64_t kw, double threshold, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                        int64_t pr, spk_sparse *out, void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_ARG(out, "Sparse output is null.");
        SPK_REQUIRE(kernel_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "Sparse conv: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(sy == 1 && sx == 1, "Stride must be 1 in sparse conv.");
        SPK_ARG(cout >= 1 && kh >= 1 && kw >= 1, "Sparse conv: kernel shape is invalid.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        const int64_t ho = in->h + pt + pb - kh + 1, wo = in->w + pl + pr - kw + 1;
        SPK_ARG(ho >= 1 && wo >= 1, "Input and kernel shapes are not compatible.");
        SPK_ARG(cout <= 255 && ho <= 255 && wo <= 255, "Sparse tensors hold at most 255 channels, rows and columns.");
        const int64_t lds = int64_t(CONV_CC) * in->c * kh * kw * 4;
        SPK_ARG(lds <= 53248, "Sparse conv: Cin*Kh*Kw can't exceed 416.");
        hipStream_t s = S(stream);
        SPK_ARG(in->t <= CONV_MAX_T, "Sparse conv: at most " << CONV_MAX_T << " steps.");
        const int64_t tiles_x = cdiv(wo, CONV_TILE), tiles = cdiv(ho, CONV_TILE) * tiles_x;
        Scratch times(std::max<int64_t>(in->n * cout * ho * wo, 1) * 2, s);
        if (in->n > 0)
        {
            const SparseConvGeom g{int32_t(in->t), int32_t(in->c), int32_t(in->h), int32_t(in->w), int32_t(cout),
                                   int32_t(kh), int32_t(kw), int32_t(pt), int32_t(pl), int32_t(ho), int32_t(wo),
                                   int32_t(tiles_x), float(threshold)};
            sparse_conv_tile<<<dim3(unsigned(tiles), unsigned(in->n), unsigned(cdiv(cout, CONV_CC))), 256,
                               size_t(lds), s>>>(in->offsets, in->events, (const float *)kernel, g,
                                                 times.as<uint16_t>());
            SPK_LAUNCHED();
        }
        build_lists(times.as<uint16_t>(), in->n, in->t, cout, ho, wo, out, s);
    });
}

int32_t spk_sparse_convstdp(const spk_sparse *in, void *kernel, int32_t kernel_type, int64_t cout, int64_t kh,
                            int64_t kw, const spk_stdp_config *configs, int64_t nconfigs, const spk_winner *winners,
                            const int64_t *counts, int64_t maxcount, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                            void *stream)
{
    return guard([&] {
        check_sparse(in);
        SPK_REQUIRE(kernel_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "STDP: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(in->n > 0, "Winners can't be empty.");
        SPK_ARG(cout >= 1 && kh >= 1 && kw >= 1, "STDP: kernel shape is invalid.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        SPK_ARG(maxcount >= 1, "STDP: winner capacity must be positive.");
        const StdpConfigs cfg = pack_configs(configs, nconfigs);
        const int64_t ho = in->h + pt + pb - kh + 1, wo = in->w + pl + pr - kw + 1;
        StdpGeom g{int32_t(in->t), int32_t(in->c), int32_t(in->h), int32_t(in->w), int32_t(cout), int32_t(kh),
                   int32_t(kw), int32_t(ho), int32_t(wo), 1, 1, int32_t(pt), int32_t(pl), in->n, maxcount,
                   int32_t(nconfigs)};
        hipStream_t s = S(stream);
        const int64_t slots = in->n * maxcount, words = cdiv(in->c * kh * kw, 32);
        SPK_ARG(slots < (1LL << 31), "STDP: too many winner slots.");
        Scratch last(std::max<int64_t>(in->n * in->c * in->h * in->w, 1) * 4, s), bits(slots * words * 4, s),
            slot(slots * 8, s), flag(4, s);
        step_map_launch(in, last.as<int32_t>(), 1, s);
        clear_flag<<<1, 1, 0, s>>>(flag.as<int32_t>());
        sparse_records<<<dim3(unsigned(cdiv(slots, 4))), 256, 0, s>>>(last.as<int32_t>(), g, winners, counts,
                                                                       bits.as<uint32_t>(), slot.as<int2>(),
                                                                       flag.as<int32_t>());
        SPK_LAUNCHED();
        stdp_apply(kernel, SPK_F32, g, cfg, slot.as<int2>(), bits.as<uint32_t>(), flag.as<int32_t>(), false, true, s);
        int32_t f = 0;
        SPK_HIP(hipMemcpyAsync(&f, flag.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SPK_HIP(hipStreamSynchronize(s));
        SPK_REQUIRE(f != 1, SPK_ERR_CONFIG, "STDP configuration index is out of range.");
        SPK_REQUIRE(f == 0, SPK_ERR_ARGUMENT, "Winner index is out of range.");
    });
}
}
