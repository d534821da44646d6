// Device pieces shared by the fused layer kernels (fused.hip, region.hip):
// FP16 fragment types, the per-pixel summary, U8 -> FP16 operands, the
// per-step fire/maximum update, the MFMA chain, and the STDP records/chains.
#pragma once

#include <cmath>
#include <vector>

#include "stdp.hpp"

namespace spk
{
namespace
{
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int MAX_STEPS = 16;  // K-steps of 32 -> K <= 512
constexpr int MAX_WAVES = 8;   // Cout <= 256
constexpr int COLS = 16;

struct Summary
{
    int32_t c;
    uint32_t mask;
    float first, max;
};

// [row tile][K step][lane][8] FP16 fragments of the FP32 masters [cout][cin][kh][kw]:
// lane l of K step s holds output channel rt*16 + l%16 and K group 4s + l/16
// (8 input channels c8*8.. of one tap); groups past the real K are zero.
// With `paired`, K-steps 2m and 2m+1 (below an odd last one) hold the groups
// 2u and 2u+1 of unit u = 4m + l/16 instead, the order in which the whole-
// sample kernel reads its 16-byte units.
__global__ void prep_fragments(const float *__restrict__ k, half8 *__restrict__ frag, int32_t cout, int32_t cin,
                               int32_t kh, int32_t kw, int32_t cin_pad, int32_t groups, int32_t steps,
                               int32_t row_tiles, int32_t paired)
{
    const int32_t total = row_tiles * steps * 64;
    for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x)
    {
        const int32_t lane = e % 64, s = (e / 64) % steps, rt = e / (64 * steps);
        const int32_t o = rt * 16 + lane % 16;
        const int32_t grp = paired && s < 2 * (steps / 2) ? 2 * (4 * (s / 2) + lane / 16) + (s & 1) : 4 * s + lane / 16;
        const int32_t cg = cin_pad / 8, tap = grp / cg, c8 = grp % cg;
        half8 v;
        for (int j = 0; j < 8; ++j)
        {
            const int32_t c = c8 * 8 + j;
            float x = 0.0f;
            if (o < cout && grp < groups && c < cin) x = k[((int64_t(o) * cin + c) * kh + tap / kw) * kw + tap % kw];
            v[j] = _Float16(x);
        }
        frag[e] = v;
    }
}

// Spike bytes -> FP16 MFMA operands: byte b zero-extended to 16 bits is the
// FP16 subnormal b * 2^-24.  The MFMA takes subnormal operands unflushed, so
// a product with an FP16 weight is exactly 2^-24 times the product with b as
// an FP16 integer, and a conv sum comes out as 2^-24 times the potential (to
// within FP32 rounding, identically for identical operands); kernels compare
// against the threshold times 2^-24 and scale their outputs back by 2^24.
__device__ __forceinline__ uint4 subnormals(uint2 raw)
{
    return make_uint4(__builtin_amdgcn_perm(0u, raw.x, 0x04010400u), __builtin_amdgcn_perm(0u, raw.x, 0x04030402u),
                      __builtin_amdgcn_perm(0u, raw.y, 0x04010400u), __builtin_amdgcn_perm(0u, raw.y, 0x04030402u));
}

constexpr float TWO_POW_24 = 16777216.0f;

// The threshold (as FP32) times 2^-24 is exact: 0, or far from underflow.
inline bool threshold_scales(double th)
{
    const float f = float(th);
    return f == 0.0f || double(f * (1.0f / TWO_POW_24)) * double(TWO_POW_24) == double(f);
}

__device__ __forceinline__ void combine(float &v, int32_t &i, float ov, int32_t oi)
{
    const bool take = (ov > v) | ((ov == v) & (oi < i));
    v = take ? ov : v;
    i = take ? oi : i;
}

// For the 4 bytes of v seen at step tt: bytes of res whose input is nonzero
// for the first time (not yet in `seen`) become tt.
__device__ __forceinline__ void first_nonzero(uint32_t &res, uint32_t &seen, uint32_t v, int32_t tt)
{
    const uint32_t nz = (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;  // 0x80 in nonzero bytes
    const uint32_t fresh = ((nz & ~seen) >> 7) * 0xFFu;
    seen |= nz;
    res = (res & ~fresh) | (uint32_t(tt) * 0x01010101u & fresh);
}

// One (channel, pixel) value of a step: shift "v > th" into its mask (newest
// step in bit 0) and keep the running maximum as a signed compare of the float
// bits, which orders every positive value correctly; the maximum is only read
// when a value fired, and then it is the largest fired value (> th >= 0).
// th - v has its sign bit set exactly when v > th (an exact difference never
// rounds across zero, and x - x is +0), and v_alignbit shifts that bit in:
// two VALU operations with no SGPR in between, so consecutive values
// pipeline instead of waiting on a compare result.
__device__ __forceinline__ void value_update(float v, float th, uint32_t &mask, int32_t &mxb)
{
    mask = __builtin_amdgcn_alignbit(mask, __builtin_bit_cast(uint32_t, th - v), 31);
    mxb = max(mxb, __builtin_bit_cast(int32_t, v));
}

// One step of one lane's 8 (channel, pixel) values (two 16-row tiles).
__device__ __forceinline__ void step_update(const float4v &acc0, const float4v &acc1, float th,
                                            uint32_t (&mask)[2][4], int32_t (&mxb)[2][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) value_update(acc0[i], th, mask[0][i], mxb[0][i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) value_update(acc1[i], th, mask[1][i], mxb[1][i]);
}

// step_update that also keeps each value at its first firing step (fv: the
// mask was 0 before it; later steps leave it).
__device__ __forceinline__ void step_update_first(const float4v &acc0, const float4v &acc1, float th,
                                                  uint32_t (&mask)[2][4], int32_t (&mxb)[2][4], float (&fv)[2][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        value_update(acc0[i], th, mask[0][i], mxb[0][i]);
        fv[0][i] = mask[0][i] <= 1u ? acc0[i] : fv[0][i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        value_update(acc1[i], th, mask[1][i], mxb[1][i]);
        fv[1][i] = mask[1][i] <= 1u ? acc1[i] : fv[1][i];
    }
}

template <int S>
__device__ __forceinline__ void load_b(half8 (&b)[S], const uint8_t *base, const int32_t (&offs)[S])
{
#pragma unroll
    for (int s = 0; s < S; ++s) b[s] = *reinterpret_cast<const half8 *>(base + offs[s]);
}

template <int S>
__device__ __forceinline__ void mfma_chain(const half8 (&a)[2][S], const half8 (&b)[S], float4v &c0, float4v &c1)
{
    c0 = float4v{0, 0, 0, 0}, c1 = float4v{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0][s], b[s], c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1][s], b[s], c1, 0, 0, 0);
    }
}


struct FusedStdp
{
    int64_t slots;  // n * count
    int32_t count, cin, cin_pad, h, w, kh, kw, pt, pl, wo, hw, t, words;
};

// Potentiate/depress bits of every winner slot, one wave per slot: bit i of
// slot e is isum <= osum for weight i = (ch, ky, kx) of the winner's kernel
// (stdp.cpp:88-118).  A winner is always its pixel's kept channel, so osum
// is the first set bit of that pixel's fire mask; isum comes from the
// first-spike map ([n][h*w][cin_pad], t outside the input).
__global__ void __launch_bounds__(256)
    fused_records(const uint8_t *__restrict__ fs, const Summary *__restrict__ summary,
                  const int32_t *__restrict__ wkey, FusedStdp p, uint32_t *__restrict__ bits)
{
    const int64_t e = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (e >= p.slots) return;
    const int32_t k = wkey[e];
    if (k < 0) return;
    const int lane = threadIdx.x & 63;
    const int64_t s = e / p.count;
    const int32_t pos = k & 0xFFFF, y = pos / p.wo - p.pt, x = pos % p.wo - p.pl;
    const uint32_t mask = summary[s * p.hw + pos].mask;
    const int32_t osum = mask ? int32_t(__builtin_ctz(mask)) : p.t;
    const int32_t taps = p.kh * p.kw, ksz = p.cin * taps;
    const uint8_t *fsn = fs + s * p.h * p.w * p.cin_pad;
    uint32_t *rec = bits + e * p.words;
    for (int32_t base = 0; base < ksz; base += 64)
    {
        const int32_t i = base + lane;
        bool bit = false;
        if (i < ksz)
        {
            const int32_t ch = i / taps, iy = y + (i / p.kw) % p.kh, ix = x + i % p.kw;
            const int32_t isum =
                (iy >= 0 && iy < p.h && ix >= 0 && ix < p.w) ? fsn[(iy * p.w + ix) * p.cin_pad + ch] : p.t;
            bit = isum <= osum;
        }
        const uint64_t b = __ballot(bit);
        if (lane == 0 && base / 32 < p.words) rec[base / 32] = uint32_t(b);
        if (lane == 32 && base / 32 + 1 < p.words) rec[base / 32 + 1] = uint32_t(b >> 32);
    }
}

// How the STDP clamp max(lower, min(v, upper)) may be evaluated.  The
// reference's selects (v < upper ? v : upper, then lower < hi ? hi : lower)
// differ from v_min_f64/v_max_f64 only on ties between a +0 and a -0, so
// the min/max instructions are exact when upper != 0 and lower != 0
// (CLAMP_MINMAX); with lower == +0 the max may return -0 where the select
// returns +0, which adding +0 corrects (CLAMP_MINMAX_P0, -0 + +0 = +0 and
// x + 0 = x otherwise); any other bound keeps the selects (CLAMP_SELECT).
enum ClampMode
{
    CLAMP_SELECT = 0,
    CLAMP_MINMAX = 1,
    CLAMP_MINMAX_P0 = 2,
    CLAMP_MED3_F32 = 3,
    CLAMP_MED3_F32_L0 = 4  // CLAMP_MED3_F32 with lower == +0: old - lower is old itself (also for -0)
};

// When both bounds are F32 values with lower <= upper, the clamp moves after
// the rounding to F32: rounding to nearest is monotone and leaves the bounds
// unchanged, so clamp-then-round equals round-then-clamp, and one v_med3_f32
// does the clamp.  v_med3_f32 (like v_max) orders -0 below +0, so a lower
// bound of +0 turns -0 into +0 as the min/max-then-+0 form does.
inline int clamp_mode(const spk_stdp_config &c)
{
    if (c.upper == 0.0) return CLAMP_SELECT;
    if (c.lower == 0.0 && std::signbit(c.lower)) return CLAMP_SELECT;
    if (double(float(c.lower)) == c.lower && double(float(c.upper)) == c.upper && c.lower <= c.upper)
        return c.lower == 0.0 ? CLAMP_MED3_F32_L0 : CLAMP_MED3_F32;
    return c.lower != 0.0 ? CLAMP_MINMAX : CLAMP_MINMAX_P0;
}

// The update of stdp_update with the clamp evaluated as MODE allows and the
// stabilize choice fixed at compile time.
template <int MODE, bool STAB>
__device__ __forceinline__ float stdp_step(float w, double rate, double lower, double upper)
{
    const double old = double(w);
    double value = STAB ? (MODE == CLAMP_MED3_F32_L0 ? old : old - lower) * (upper - old) : 1.0;
    value = __builtin_fma(rate, value, old);
    if (MODE == CLAMP_MED3_F32 || MODE == CLAMP_MED3_F32_L0)
        return __builtin_amdgcn_fmed3f(float(value), float(lower), float(upper));
    double r;
    if (MODE == CLAMP_SELECT)
    {
        const double hi = value < upper ? value : upper;
        r = lower < hi ? hi : lower;
    }
    else
    {
        r = __builtin_fmax(__builtin_fmin(value, upper), lower);
        if (MODE == CLAMP_MINMAX_P0) r = r + 0.0;
    }
    return float(r);
}

constexpr int STDP_THREADS = 256;
constexpr int STDP_LIST = 1024;  // winners staged in LDS per pass

// The update chains of the fused step (stdp.cpp:139-154): the winners of
// every sample in rank order, samples in order, each updating the kernel of
// its channel; all use configuration 0.  One block per (output channel,
// slice of STDP_THREADS weights): each thread scans a contiguous range of
// winner slots, one block-wide scan lists the slots naming the channel in
// slot order, their bit words are staged in LDS, and every thread runs the
// chain of its weight 32 updates at a time from a register of bits.
template <int MODE, bool STAB>
__global__ void __launch_bounds__(STDP_THREADS)
    fused_stdp(float *__restrict__ kernel, const uint32_t *__restrict__ bits, const int32_t *__restrict__ wkey,
               FusedStdp p, spk_stdp_config cfg)
{
    __shared__ int32_t list[STDP_LIST];
    __shared__ uint32_t lbits[STDP_LIST * (STDP_THREADS / 32)];  // the slice's words of each listed winner
    __shared__ int32_t wave_sum[STDP_THREADS / 64];
    const int32_t key = blockIdx.x, ksz = p.cin * p.kh * p.kw;
    const int32_t i = blockIdx.y * STDP_THREADS + threadIdx.x;
    const bool live = i < ksz;
    float *kz = kernel + int64_t(key) * ksz;
    float wr = live ? kz[i] : 0.0f;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this slice's words of a winner's record
    const int32_t word0 = blockIdx.y * (STDP_THREADS / 32);
    const int32_t sw = min(STDP_THREADS / 32, p.words - word0), my_word = (i / 32) - word0;

    // the thread's slots [e0, e1), a multiple of 4 wide so keys load as int4
    const int64_t per = (p.slots + 4 * STDP_THREADS - 1) / (4 * STDP_THREADS) * 4;
    const int64_t e0 = min(int64_t(threadIdx.x) * per, p.slots), e1 = min(e0 + per, p.slots);
    auto keys4 = [&](int64_t e) {
        if (e + 4 <= p.slots) return *reinterpret_cast<const int4 *>(wkey + e);
        int4 k = make_int4(-1, -1, -1, -1);
        if (e < p.slots) k.x = wkey[e];
        if (e + 1 < p.slots) k.y = wkey[e + 1];
        if (e + 2 < p.slots) k.z = wkey[e + 2];
        return k;
    };
    auto hit = [&](int32_t k) { return (k >> 16) == key; };
    int32_t cnt = 0;
    for (int64_t e = e0; e < e1; e += 4)
    {
        const int4 k = keys4(e);
        cnt += hit(k.x) + hit(k.y) + hit(k.z) + hit(k.w);
    }
    int32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const int32_t o = __shfl_up(incl, d, 64);
        incl += lane >= d ? o : 0;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    int32_t first = incl - cnt, m = 0;
    for (int v = 0; v < STDP_THREADS / 64; ++v)
    {
        first += v < wave ? wave_sum[v] : 0;
        m += wave_sum[v];
    }

    for (int32_t pass = 0; pass < m; pass += STDP_LIST)
    {
        // the thread's matches with list index in [pass, pass + STDP_LIST)
        if (first < pass + STDP_LIST && first + cnt > pass)
        {
            int32_t at = first;
            for (int64_t e = e0; e < e1; e += 4)
            {
                const int4 k4 = keys4(e);
                const int32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (hit(kk[u]))
                    {
                        if (at >= pass && at < pass + STDP_LIST) list[at - pass] = int32_t(e + u);
                        ++at;
                    }
            }
        }
        __syncthreads();
        const int32_t mm = min(STDP_LIST, m - pass);
        for (int32_t q = threadIdx.x; q < mm * sw; q += STDP_THREADS)
            lbits[q] = bits[int64_t(list[q / sw]) * p.words + word0 + q % sw];
        __syncthreads();
        // bits of 32 updates of this weight from the staged words
        auto gather = [&](int32_t q0) {
            uint32_t b = 0;
#pragma unroll
            for (int u = 0; u < 32; ++u) b |= ((lbits[min(q0 + u, mm - 1) * sw + my_word] >> (i & 31)) & 1u) << u;
            return b;
        };
        uint32_t b = live ? gather(0) : 0u;
        if (live)
            for (int32_t q0 = 0; q0 < mm; q0 += 32)
            {
                // the next chunk's bits are read while this chunk's chain runs
                const uint32_t b_next = q0 + 32 < mm ? gather(q0 + 32) : 0u;
                const int32_t n_up = min(32, mm - q0);
                if (n_up == 32)
                {
#pragma unroll
                    for (int u = 0; u < 32; ++u)
                        wr = stdp_step<MODE, STAB>(wr, (b >> u) & 1u ? cfg.positive : cfg.negative, cfg.lower,
                                                   cfg.upper);
                }
                else
                    for (int u = 0; u < n_up; ++u)
                        wr = stdp_step<MODE, STAB>(wr, (b >> u) & 1u ? cfg.positive : cfg.negative, cfg.lower,
                                                   cfg.upper);
                b = b_next;
            }
        __syncthreads();
    }
    if (live) kz[i] = wr;
}


// Launches fused_stdp for the clamp and stabilize choice of cfg.
inline void launch_fused_stdp(dim3 grid, hipStream_t s, float *kernel, const uint32_t *bits, const int32_t *wkey,
                              const FusedStdp &p, const spk_stdp_config &cfg)
{
    const int mode = clamp_mode(cfg);
    const bool stab = cfg.stabilize != 0;
#define STDP_CASE(M, B) \
    if (mode == M && stab == B) fused_stdp<M, B><<<grid, STDP_THREADS, 0, s>>>(kernel, bits, wkey, p, cfg);
    STDP_CASE(CLAMP_SELECT, false) STDP_CASE(CLAMP_SELECT, true) STDP_CASE(CLAMP_MINMAX, false)
    STDP_CASE(CLAMP_MINMAX, true) STDP_CASE(CLAMP_MINMAX_P0, false) STDP_CASE(CLAMP_MINMAX_P0, true)
    STDP_CASE(CLAMP_MED3_F32, false) STDP_CASE(CLAMP_MED3_F32, true) STDP_CASE(CLAMP_MED3_F32_L0, false)
    STDP_CASE(CLAMP_MED3_F32_L0, true)
#undef STDP_CASE
}

}  // namespace
}  // namespace spk

namespace spk
{
// Optional HIP-event timing of the main fused kernels (spk_kernel_timer):
// every launch records an event pair on its stream while enabled.
struct KernelTimer
{
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
KernelTimer &timer();

// Work forked from a call's stream onto the device's side stream (the
// caller's stream when `use` is false) and joined back by join(), or by the
// destructor on an error path.  Fork and join are event record/wait pairs,
// so a stream capture of the call follows them.
struct SideStream
{
    hipStream_t main, stream;
    hipEvent_t fork = nullptr, done = nullptr;
    bool joined = false;
    SideStream(hipStream_t main, bool use);
    SideStream(const SideStream &) = delete;
    SideStream &operator=(const SideStream &) = delete;
    void again();  // the side stream waits for the work the main stream has now
    void join();
    ~SideStream();
};

// The layer step over output regions (region.hip) for a uniform batch whose
// sample does not fit the LDS whole; same arguments as spk_stdp_layer_fused,
// cfg = null without learning.
// The fused layer step (spk_stdp_layer_fused).  scratch = null: its buffers
// come from the stream's workspace; otherwise from scratch[0, scratch_bytes),
// and a call that needs more, or the region path, is refused.
void stdp_layer_fused(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, void *kernel,
                      int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                      double threshold, int64_t ry, int64_t rx, int64_t count, const spk_stdp_config *configs,
                      int64_t nconfigs, void *out, void *pooled, int64_t pkh, int64_t pkw, int64_t psh, int64_t psw,
                      spk_winner *winners, int64_t *counts, void *potentials, void *stream, uint8_t *scratch,
                      int64_t scratch_bytes);

void region_layer_step(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, float *kernel,
                       int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                       double threshold, int64_t ry, int64_t rx, int64_t count, const spk_stdp_config *cfg,
                       uint8_t *out, uint8_t *pooled, int64_t pkh, int64_t pkw, int64_t psh, int64_t psw,
                       spk_winner *winners, int64_t *counts, float *pot, hipStream_t s);
}  // namespace spk
