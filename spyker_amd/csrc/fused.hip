// Fused FP16-MFMA layer step: conv -> threshold -> inhibit -> fire summary,
// then convwta, spike/pool writer and STDP on the FP32 master weights.
//
// Semantics are the composition the reference runs as separate calls
// (src/spyker/opers.cpp:408-423 conv, 472-478 threshold, 639-646 inhibit,
// 656-662 convwta, 522-552 fire, 600-637 pool, 679-693 STDP) with the
// inhibit/convwta/fire thresholds at their default 0, a conv threshold
// th >= 0 and replacement value 0.  Under those conditions every value after
// the threshold is 0 or > th >= 0, so for each sample and output pixel the
// whole post-inhibit tensor is described by
//
//     c*     the kept channel (inhibit.cpp:64-85: first channel maximum at
//            step T - max(cnt, 1), cnt = #steps whose maximum is > 0)
//     mask   bit t set iff the kept channel's value at t is > 0 (= fire)
//     first  the value at the first set bit (convwta's t0 value)
//     max    the largest value over t (for convwta's per-sample max)
//
// and every other channel is 0 there.  The conv is the dense time-unrolled
// contraction: per sample and step a [Cout x K] x [K x pixels] product with
// K = taps * Cin, on v_mfma_f32_16x16x32_f16 (spikes are exact in FP16; the
// weights are the FP32 masters rounded to FP16; accumulation is FP32).
//
// Layout of the main kernel (one workgroup of NW waves per CU, persistent
// over samples):
//   - wave w owns output channels [32w, 32w+32) as two 16-row MFMA tiles;
//     its FP16 weight fragments for all K stay in VGPRs for the whole kernel;
//   - the sample's input for all T steps is staged once in LDS as U8
//     [t][y][x][cin_pad] (channel innermost) inside a zero border of the
//     conv's padding, so every tap of every output pixel is in the stage;
//   - columns are 16 row-major output pixels.  A lane's B fragment is 8
//     bytes of the stage (8 channels at one tap of its pixel) zero-extended
//     to 16 bits: the FP16 subnormals b * 2^-24, which the MFMA multiplies
//     exactly, so the potentials come out exactly 2^-24 times the sums with
//     b as FP16 integers (the threshold is scaled to match).  Every wave
//     runs a tile's T steps without barriers, loading the bytes of step k+2
//     while the MFMAs of step k+1 run;
//   - per step and (channel, pixel) a lane only shifts "v > th" into a mask
//     and keeps the running maximum;
//   - at the end of a tile, cnt = popcount(OR of all masks) gives inhibit's
//     step ts = T - max(cnt, 1); one MFMA pass with each lane reading its
//     pixel's bytes at that pixel's own ts yields the kept channel c*, and a
//     second at c*'s first firing step yields convwta's t0 value.  The
//     passes issue the same MFMA sequence on the same operands as the step
//     loop, so they reproduce its potentials bit for bit.  Two barriers a
//     tile exchange the per-pixel partials between waves.
#include <array>
#include <atomic>
#include <map>
#include <mutex>

#include "fused.hpp"

namespace spk
{
namespace
{
struct FusedGeom
{
    int32_t n, t, cin, h, w, cout, kh, kw, pt, pl, ho, wo;
    int32_t ph, pw, cin_pad, groups, steps, waves, tiles;
    int32_t stage_bytes;  // bytes of the staged sample ([t][ph][pw][cin_pad] U8, 16-aligned)
    int32_t sstep;        // bytes of one staged step (ph * pw * cin_pad)
    int32_t stage_fast;   // staging by quad transposes (hw % 4 == 0, aligned input, cin_pad/4 | block quads)
    float th;             // the conv threshold times 2^-24 (potentials are computed times 2^-24)
    float thf;            // the conv threshold (the int8 kernel scales it to its fixed point)
    int32_t upr, units;   // int8 kernel: 16-byte K-units per window row, and in all
    int32_t pkh, pkw, psh, psw, pho, pwo;  // max-pool of the spikes (pho = 0: none)
    int32_t zp;                            // bytes of a sample's pooled spikes each tile zeroes
};

// Zeroes bytes [lo, hi) of dst with the block's threads: 16-byte stores on
// the aligned middle, byte stores at the two ends.
// a 16-byte zero store marked non-temporal: the zeroed output is not read
// again here, and streaming it past the L2 keeps the cache for the lines the
// 1s are scattered into (measured 3 % off the layer step)
__device__ __forceinline__ void zero16(void *p)
{
    using u32x4 = __attribute__((ext_vector_type(4))) uint32_t;
    __builtin_nontemporal_store(u32x4{0u, 0u, 0u, 0u}, reinterpret_cast<u32x4 *>(p));
}

__device__ __forceinline__ void zero_slice(uint8_t *dst, int64_t lo, int64_t hi)
{
    if (lo >= hi) return;
    int64_t head = int64_t((16 - (reinterpret_cast<uintptr_t>(dst + lo) & 15)) & 15);
    head = head < hi - lo ? head : hi - lo;
    if (int64_t(threadIdx.x) < head) dst[lo + threadIdx.x] = 0;
    const int64_t body = lo + head, units = (hi - body) / 16;
    uint4 *b = reinterpret_cast<uint4 *>(dst + body);
    for (int64_t i = threadIdx.x; i < units; i += blockDim.x) b[i] = make_uint4(0, 0, 0, 0);
    for (int64_t i = body + units * 16 + threadIdx.x; i < hi; i += blockDim.x) dst[i] = 0;
}

// Byte offset of K-group grp (8 input channels at one tap) in the stage,
// relative to the output pixel's own position (the top-left tap of its
// window in the zero-bordered stage).  Groups past the real K (zero weights)
// reuse group 0's operands.
__device__ __forceinline__ int32_t tap_off(const FusedGeom &g, int32_t grp)
{
    const int32_t cg = g.cin_pad / 8;
    const int32_t gg = grp < g.groups ? grp : 0, tap = gg / cg, c8 = gg % cg;
    return ((tap / g.kw) * g.pw + tap % g.kw) * g.cin_pad + c8 * 8;
}

// The bytes of one step's B fragments for a lane.  K-step s takes K-group
// 4s+q (q = lane / 16) by default.  PAIRED (channels padded to a multiple of
// 16, so 16-byte units never straddle a tap) lets K-steps 2m and 2m+1 share
// one 16-byte read: unit u = 4m+q, whose low 8 bytes are K-step 2m's group 2u
// and high 8 bytes K-step 2m+1's group 2u+1; an odd last K-step reads its 8
// bytes alone.  prep_fragments lays the weights out in the same order.
template <int S, bool PAIRED>
struct StepBytes
{
    static constexpr int P = PAIRED ? S / 2 : 0, L = S - 2 * P;
    uint4 p[P > 0 ? P : 1];
    uint2 l[L > 0 ? L : 1];
    __device__ __forceinline__ uint2 frag(int s) const
    {
        if (s < 2 * P) return (s & 1) ? make_uint2(p[s / 2].z, p[s / 2].w) : make_uint2(p[s / 2].x, p[s / 2].y);
        return l[s - 2 * P];
    }
};

// Stage offsets of a lane's reads relative to its pixel (see StepBytes).
template <int S, bool PAIRED>
struct StepOffsets
{
    static constexpr int P = StepBytes<S, PAIRED>::P, L = StepBytes<S, PAIRED>::L;
    int32_t p[P > 0 ? P : 1], l[L > 0 ? L : 1];
    __device__ __forceinline__ StepOffsets(const FusedGeom &g, int32_t q)
    {
#pragma unroll
        for (int i = 0; i < P; ++i) p[i] = tap_off(g, 2 * (4 * i + q));
#pragma unroll
        for (int i = 0; i < L; ++i) l[i] = tap_off(g, 4 * (2 * P + i) + q);
    }
};

template <int S, bool PAIRED>
__device__ __forceinline__ void load_raw(StepBytes<S, PAIRED> &r, const uint8_t *base,
                                         const StepOffsets<S, PAIRED> &o)
{
#pragma unroll
    for (int i = 0; i < StepBytes<S, PAIRED>::P; ++i) r.p[i] = *reinterpret_cast<const uint4 *>(base + o.p[i]);
#pragma unroll
    for (int i = 0; i < StepBytes<S, PAIRED>::L; ++i) r.l[i] = *reinterpret_cast<const uint2 *>(base + o.l[i]);
}

// The MFMA chain of one step on raw spike bytes (as FP16 subnormals), for
// the RT 16-row tiles of a wave.
template <int RT, int S, bool PAIRED>
__device__ __forceinline__ void mfma_sub(const half8 (&a)[RT][S], const StepBytes<S, PAIRED> &r, float4v (&c)[RT])
{
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) c[rt] = float4v{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        const half8 b = __builtin_bit_cast(half8, subnormals(r.frag(s)));
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) c[rt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[rt][s], b, c[rt], 0, 0, 0);
    }
}

// The update of every value of one step (see value_update).
template <int RT>
__device__ __forceinline__ void update_all(const float4v (&u)[RT], float th, uint32_t (&mask)[RT][4],
                                           int32_t (&mxb)[RT][4])
{
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int i = 0; i < 4; ++i) value_update(u[rt][i], th, mask[rt][i], mxb[rt][i]);
}

// One step of a tile, interleaved for a single wave: the MFMA chain of step
// k+1 on its raw bytes `rin`, and between its MFMAs the update of step k's
// values `u` and the loads of step k+2's bytes into `rout`.  The conversions,
// value updates and addresses ride along with the MFMAs, so one wave keeps
// the matrix core and the VALU busy together instead of alternating between
// an MFMA phase and a VALU phase.  Value v is updated after MFMA number
// RT + v * (S*RT - RT) / (4*RT): the first fragment's MFMAs give the previous
// chain's last results time to land.
template <int RT, int S, bool PAIRED>
__device__ __forceinline__ void step_iter(const half8 (&a)[RT][S], const StepBytes<S, PAIRED> &rin, float4v (&c)[RT],
                                          const float4v (&u)[RT], float th, uint32_t (&mask)[RT][4],
                                          int32_t (&mxb)[RT][4], StepBytes<S, PAIRED> &rout, const uint8_t *next,
                                          const StepOffsets<S, PAIRED> &o)
{
    constexpr int V = 4 * RT, SPAN = S * RT - RT, P = StepBytes<S, PAIRED>::P;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) c[rt] = float4v{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        const half8 b = __builtin_bit_cast(half8, subnormals(rin.frag(s)));
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
        {
            c[rt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[rt][s], b, c[rt], 0, 0, 0);
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (SPAN > 0 && RT + (v * SPAN) / V == s * RT + rt)
                    value_update(u[v / 4][v % 4], th, mask[v / 4][v % 4], mxb[v / 4][v % 4]);
        }
        // the read that fed this fragment, for the step after next
        if (s < 2 * P && (s & 1)) rout.p[s / 2] = *reinterpret_cast<const uint4 *>(next + o.p[s / 2]);
        if (s >= 2 * P) rout.l[s - 2 * P] = *reinterpret_cast<const uint2 *>(next + o.l[s - 2 * P]);
    }
    if (SPAN <= 0) update_all(u, th, mask, mxb);
    // keep the next iteration's conversions below this one's loads
    __builtin_amdgcn_sched_barrier(0);
}

// RT 16-row tiles per wave: 4 (64 channels, one wave per SIMD, with the
// accumulation registers of the whole file) halves the byte conversions and
// LDS reads per MFMA against 2 (32 channels, two waves per SIMD).
template <int RT>
constexpr int layer_threads()
{
    return RT == 4 ? 256 : 512;
}

// Stages sample n of the [n][t][cin][h][w] input as [t][y][x][cin_pad] bytes
// inside the zero border of the conv's padding (written once per kernel).
// Every staged byte is XORed with each byte of `bias` (0, or 0x80 for the
// int8 kernel, whose border holds 0x80 too).
__device__ __forceinline__ void stage_sample(const uint8_t *__restrict__ in, int32_t n, const FusedGeom &g,
                                             uint8_t *stage, uint32_t bias)
{
    const int32_t hwi = g.h * g.w;
    if (g.stage_fast)
    {
        // a lane loads one dword (4 pixels of channel c) of the
        // [t][cin][h][w] input; the 4 lanes of a quad hold channels
        // 4*c4 .. 4*c4+3 of the same 4 pixels and transpose the 4x4 bytes
        // (two DPP exchanges + v_perm), after which lane l4 holds the 4
        // channels of pixel l4: one dword store per lane
        const int32_t D = g.cin_pad / 4, npq = hwi / 4, step = int32_t(blockDim.x) / 4 / D;
        const int32_t Q = int32_t(threadIdx.x) >> 2, l4 = int32_t(threadIdx.x) & 3, c4 = Q % D;
        const int32_t c = 4 * c4 + l4;
        const uint32_t sel1 = (l4 & 1) ? 0x03070105u : 0x06020400u;
        const uint32_t sel2 = (l4 & 2) ? 0x03020706u : 0x05040100u;
        // raw buffer loads of the sample: 32-bit offsets, and offsets past
        // its end (the walk's tail) read 0
        const __amdgpu_buffer_rsrc_t src = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint8_t *>(in + int64_t(n) * g.t * g.cin * hwi), 0, g.t * g.cin * hwi, 0x00020000);
        uint32_t *dst = reinterpret_cast<uint32_t *>(stage);
        // lin = tt * npq + pq walks the (step, pixel quad) pairs by `step`
        // (<= npq, so pq wraps at most once per move); off is the byte
        // offset of the lane's dword, channel c of step tt.  U loads are
        // in flight per lane before the first store.
        int32_t lin = Q / D, pq = lin, off = (c * npq + pq) * 4;
        const int32_t total = g.t * npq, wrap_add = (g.cin - 1) * npq * 4;
        // the store side walks the same pairs as (tt, pixel row, column)
        // of the lane's own pixel 4*pq + l4, moving 4*step pixels a time:
        // dy rows and dx columns, one more row when the column wraps, and
        // back h rows into the next step when the pixel passes h*w
        int32_t sq = pq, st = 0, sy = (4 * pq + l4) / g.w, sx = (4 * pq + l4) % g.w;
        const int32_t dy = 4 * step / g.w, dx = 4 * step % g.w;
        constexpr int U = 16;
        for (; lin < total; lin += U * step)
        {
            uint32_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                v[u] = uint32_t(__builtin_amdgcn_raw_buffer_load_b32(src, off, 0, 0));
                pq += step, off += step * 4;
                const bool wrap = pq >= npq;
                pq = wrap ? pq - npq : pq;
                off = wrap ? off + wrap_add : off;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const uint32_t x = c < g.cin ? v[u] : 0u;
                const uint32_t y = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(x), 0xB1, 0xF, 0xF, true));  // lane^1
                const uint32_t r1 = __builtin_amdgcn_perm(y, x, sel1);
                const uint32_t y2 = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(r1), 0x4E, 0xF, 0xF, true));  // lane^2
                if (lin + u * step < total)
                    dst[((st * g.ph + sy + g.pt) * g.pw + sx + g.pl) * D + c4] =
                        __builtin_amdgcn_perm(y2, r1, sel2) ^ bias;
                sq += step, sy += dy, sx += dx;
                sy = sx >= g.w ? sy + 1 : sy;
                sx = sx >= g.w ? sx - g.w : sx;
                const bool wrap = sq >= npq;
                sq = wrap ? sq - npq : sq;
                st = wrap ? st + 1 : st;
                sy = wrap ? sy - g.h : sy;
            }
        }
    }
    else
    {
        // any other shape: scatter the [t][cin][h][w] input byte by byte
        // (the padded channels keep the border's value)
        const uint8_t *src = in + int64_t(n) * g.t * g.cin * hwi;
        const int32_t chw = g.cin * hwi, total = g.t * chw;
        for (int32_t i = int32_t(threadIdx.x); i < total; i += int32_t(blockDim.x))
        {
            const int32_t tt = i / chw, r = i - tt * chw, c = r / hwi, pix = r - c * hwi, yy = pix / g.w;
            const int32_t xx = pix - yy * g.w;
            stage[((tt * g.ph + yy + g.pt) * g.pw + xx + g.pl) * g.cin_pad + c] = src[i] ^ uint8_t(bias);
        }
    }
}

// The first spiking step of every staged input byte of sample n (t when it
// never spikes), 8 channels of one pixel per thread: [n][h*w][cin_pad]; the
// stage's bytes are the input's XORed with `bias`.
__device__ __forceinline__ void first_spike_map(const uint8_t *stage, int32_t n, const FusedGeom &g,
                                                uint8_t *__restrict__ fs_out, uint32_t bias)
{
    const int32_t hwi = g.h * g.w, cg = g.cin_pad / 8;
    const uint32_t tb = uint32_t(g.t) * 0x01010101u;
    for (int32_t u = threadIdx.x; u < hwi * cg; u += blockDim.x)
    {
        const int32_t pix = u / cg, c8 = u - pix * cg, py = pix / g.w, px = pix - py * g.w;
        const uint8_t *sp = stage + ((py + g.pt) * g.pw + px + g.pl) * g.cin_pad + c8 * 8;
        uint32_t lo = tb, hi = tb, seen_lo = 0, seen_hi = 0;
        for (int32_t tt = 0; tt < g.t; ++tt)
        {
            const uint2 v = *reinterpret_cast<const uint2 *>(sp + tt * g.sstep);
            first_nonzero(lo, seen_lo, v.x ^ bias, tt);
            first_nonzero(hi, seen_hi, v.y ^ bias, tt);
        }
        *reinterpret_cast<uint2 *>(fs_out + (int64_t(n) * hwi * cg + u) * 8) = make_uint2(lo, hi);
    }
}

// each sample's 1s: its kept channel at every step it fires, in its
// spike plane and in every pool window that holds the pixel (the max of a
// window of 0/1 bytes is 1 exactly when one of its pixels is; pixels of a
// window store the same value).  One (step, pixel) pair a thread, from
// smap; the sample's zeros were stored before the last barrier.
__device__ __forceinline__ void scatter_sample(int32_t m, const uint2 *smap, uint8_t *__restrict__ out,
                                               uint8_t *__restrict__ pooled, const FusedGeom &g)
{
    const int32_t hw = g.ho * g.wo;
    const int64_t s_bytes = int64_t(g.t) * g.cout * hw, p_bytes = int64_t(g.t) * g.cout * g.pho * g.pwo;
    uint8_t *s_out = out ? out + int64_t(m) * s_bytes : nullptr;
    uint8_t *p_out = pooled ? pooled + int64_t(m) * p_bytes : nullptr;
    int32_t st = int32_t(threadIdx.x) / hw, q2 = int32_t(threadIdx.x) - st * hw;
    const int32_t bst = int32_t(blockDim.x) / hw, bq = int32_t(blockDim.x) - bst * hw;
    for (; st < g.t; st += bst, q2 += bq)
    {
        if (q2 >= hw) q2 -= hw, ++st;
        if (st >= g.t) break;
        const uint2 e = smap[q2];
        if (!((e.y >> st) & 1u)) continue;
        if (s_out) s_out[(int64_t(st) * g.cout + e.x) * hw + q2] = 1;
        if (p_out)
        {
            const int32_t y = q2 / g.wo, x = q2 % g.wo;
            const int32_t py0 = y >= g.pkh ? (y - g.pkh) / g.psh + 1 : 0, py1 = min(y / max(g.psh, 1), g.pho - 1);
            const int32_t px0 = x >= g.pkw ? (x - g.pkw) / g.psw + 1 : 0, px1 = min(x / max(g.psw, 1), g.pwo - 1);
            for (int32_t py = py0; py <= py1; ++py)
                for (int32_t px = px0; px <= px1; ++px)
                    p_out[((int64_t(st) * g.cout + e.x) * g.pho + py) * g.pwo + px] = 1;
        }
    }
}

template <int RT, int S, bool PAIRED, bool DIAG>
__global__ void __launch_bounds__(layer_threads<RT>())
    fused_layer(const uint8_t *__restrict__ in, const half8 *__restrict__ frag, Summary *__restrict__ summary,
                float *__restrict__ pot_out, uint8_t *__restrict__ fs_out, uint8_t *__restrict__ out,
                uint8_t *__restrict__ pooled, FusedGeom g)
{
    extern __shared__ __align__(16) uint8_t lds[];
    uint8_t *stage = lds;                                                       // [t][ph][pw][cin_pad] U8
    uint32_t *any_s = reinterpret_cast<uint32_t *>(lds + g.stage_bytes);        // [2][MAX_WAVES][COLS]
    float *red_v = reinterpret_cast<float *>(any_s + 2 * MAX_WAVES * COLS);     // [2][MAX_WAVES][COLS]
    int32_t *red_i = reinterpret_cast<int32_t *>(red_v + 2 * MAX_WAVES * COLS);  // [2][MAX_WAVES][COLS]
    uint2 *smap = reinterpret_cast<uint2 *>(red_i + 2 * MAX_WAVES * COLS);       // [ho*wo] {c*, fire mask}

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
    const int32_t hw = g.ho * g.wo, hwi = g.h * g.w, cg = g.cin_pad / 8;
    const int32_t o0 = w * 16 * RT + 4 * q;  // the lane's first channel; value (r, i) is channel o0 + 16r + i

    // weight fragments and tap offsets, resident for the whole kernel
    half8 a[RT][S];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int r = 0; r < RT; ++r) a[r][s] = frag[((w * RT + r) * S + s) * 64 + lane];
    // wait for the fragments here: with their loads still pending where the
    // tile loop (global stores, no global loads) is first entered, the
    // compiler drains every outstanding store before each sample's tiles
    __builtin_amdgcn_s_waitcnt(0);
    const StepOffsets<S, PAIRED> toff(g, q);
    // the zero border and the padded channels are written once; every sample
    // rewrites only the interior
    for (int32_t i = threadIdx.x; i < g.stage_bytes / 16; i += blockDim.x)
        reinterpret_cast<uint4 *>(stage)[i] = make_uint4(0, 0, 0, 0);

    int32_t prev = -1;  // the sample whose 1s are not stored yet
    for (int32_t n = blockIdx.x; n < g.n; n += gridDim.x)
    {
        __syncthreads();
        stage_sample(in, n, g, stage, 0u);
        __syncthreads();
        if (fs_out) first_spike_map(stage, n, g, fs_out, 0u);
        // the previous sample's 1s, after this sample's staging so that its
        // loads do not wait behind these stores; smap still holds that
        // sample's map (the first tile's end rewrites it, after two barriers)
        if (prev >= 0 && (out || pooled)) scatter_sample(prev, smap, out, pooled, g);

        // the sample's spike and pooled blocks: all 0 but the kept channels'
        // 1s.  The aligned 16-byte units of the spike block are zeroed a few
        // at a time inside the step loops (so the stores stream out while the
        // MFMAs run instead of bursting at once on every CU), the pooled
        // block a slice a tile, the block's edge bytes after the last tile;
        // the 1s are scattered after the next barriers (at the next sample's
        // start, or after the last sample).
        const int64_t s_bytes = int64_t(g.t) * g.cout * hw, p_bytes = int64_t(g.t) * g.cout * g.pho * g.pwo;
        uint8_t *s_out = out ? out + int64_t(n) * s_bytes : nullptr;
        uint8_t *p_out = pooled ? pooled + int64_t(n) * p_bytes : nullptr;
        int64_t s_head = 0, s_units = 0, s_upt = 0;  // edge bytes, 16-byte units, units a tile
        if (s_out)
        {
            s_head = min(int64_t((16 - (reinterpret_cast<uintptr_t>(s_out) & 15)) & 15), s_bytes);
            s_units = (s_bytes - s_head) / 16;
            s_upt = (s_units + g.tiles - 1) / g.tiles;
        }
        uint4 *s_body = reinterpret_cast<uint4 *>(s_out + s_head);
        for (int32_t tile = 0; tile < g.tiles; ++tile)
        {
            if (p_out) zero_slice(p_out, int64_t(tile) * g.zp, min(int64_t(tile + 1) * g.zp, p_bytes));
            const int64_t u0 = int64_t(tile) * s_upt, u1 = min(u0 + s_upt, s_units);
            // reduction buffers alternate between tiles, so two barriers a
            // tile order every write after the reads of the same buffer
            const int32_t buf = (tile & 1) * MAX_WAVES * COLS;
            const int32_t p = tile * COLS + j;
            const bool pvalid = p < hw;
            const int32_t pc = pvalid ? p : hw - 1, y = pc / g.wo, x = pc % g.wo;
            const uint8_t *pb = stage + (y * g.pw + x) * g.cin_pad;  // the pixel's window at step 0
            uint32_t mask[RT][4];
            int32_t mxb[RT][4];
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) mask[r][i] = 0, mxb[r][i] = 0;

            // the steps, no barriers: each iteration runs the MFMAs of step
            // k+1 with the update of step k and the loads of step k+2 (loads
            // past the last step re-read it, unused)
            {
                const int32_t last = g.t - 1;
                StepBytes<S, PAIRED> rA, rB;
                float4v cA[RT], cB[RT];
                auto diag = [&](int32_t k, const float4v (&u)[RT]) {
                    if (DIAG && pvalid)
#pragma unroll
                        for (int r = 0; r < RT; ++r)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                            {
                                const int32_t o = o0 + r * 16 + i;
                                if (o < g.cout)
                                    pot_out[((int64_t(n) * g.t + k) * g.cout + o) * hw + p] = u[r][i] * TWO_POW_24;
                            }
                };
                load_raw(rA, pb, toff);
                load_raw(rB, pb + min(1, last) * g.sstep, toff);
                mfma_sub(a, rA, cA);
                // pairs of iterations with no branch between them, so the
                // loads of one stay ahead of the MFMAs of the next
                int32_t k = 0;
                int64_t zu = u0 + int32_t(threadIdx.x);  // the thread's next zero unit of this tile
                for (; k + 2 < g.t; k += 2)
                {
                    // cA: step k; rB: the bytes of step k+1
                    step_iter(a, rB, cB, cA, g.th, mask, mxb, rA, pb + (k + 2) * g.sstep, toff);
                    diag(k, cA);
                    if (zu < u1) zero16(s_body + zu);
                    zu += blockDim.x;
                    step_iter(a, rA, cA, cB, g.th, mask, mxb, rB, pb + min(k + 3, last) * g.sstep, toff);
                    diag(k + 1, cB);
                }
                if (k + 1 < g.t)
                {
                    step_iter(a, rB, cB, cA, g.th, mask, mxb, rA, pb + last * g.sstep, toff);
                    diag(k, cA);
                    update_all(cB, g.th, mask, mxb);
                    diag(k + 1, cB);
                }
                else
                {
                    update_all(cA, g.th, mask, mxb);
                    diag(k, cA);
                }
                for (; zu < u1; zu += blockDim.x) zero16(s_body + zu);
            }

            // steps at which some channel fired, per pixel (bit T-1-t for step t)
            {
                uint32_t any = 0;
#pragma unroll
                for (int r = 0; r < RT; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) any |= mask[r][i];
                any |= __shfl_xor(any, 16, 64);
                any |= __shfl_xor(any, 32, 64);
                if (lane < COLS) any_s[buf + w * COLS + lane] = any;
            }
            __syncthreads();
            // pass 1: the potentials at inhibit's step ts = T - max(cnt, 1),
            // each lane reading its pixel's operands at that pixel's own ts
            // (the same bytes the step loop read there, so the same sums); the
            // kept channel is the first maximum of the thresholded values
            // (inhibit.cpp:64-85).  The lane's channels ascend with (r, i), so
            // strict '>' keeps the first maximum; padded channels have zero
            // weights and value 0, which never beats a lower channel.
            float4v gv[RT];
            int32_t ti;  // inhibit's step at this pixel
            {
                uint32_t all = 0;
                for (int32_t ow = 0; ow < g.waves; ++ow) all |= any_s[buf + ow * COLS + j];
                const int32_t cnt = __builtin_popcount(all);
                ti = g.t - (cnt > 1 ? cnt : 1);
                StepBytes<S, PAIRED> r;
                load_raw(r, pb + ti * g.sstep, toff);
                mfma_sub(a, r, gv);
                float bv = -1.0f;
                int32_t bi = INT32_MAX;
#pragma unroll
                for (int r2 = 0; r2 < RT; ++r2)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                    {
                        const float v = gv[r2][i];
                        const float xp = v > g.th ? v : 0.0f;
                        const bool take = xp > bv;
                        bv = take ? xp : bv;
                        bi = take ? o0 + r2 * 16 + i : bi;
                    }
                combine(bv, bi, __shfl_xor(bv, 16, 64), __shfl_xor(bi, 16, 64));
                combine(bv, bi, __shfl_xor(bv, 32, 64), __shfl_xor(bi, 32, 64));
                if (lane < COLS) red_v[buf + w * COLS + lane] = bv, red_i[buf + w * COLS + lane] = bi;
            }
            __syncthreads();
            int32_t cs;
            {
                float m = red_v[buf + j];
                cs = red_i[buf + j];
                for (int32_t ow = 1; ow < g.waves; ++ow)
                    combine(m, cs, red_v[buf + ow * COLS + j], red_i[buf + ow * COLS + j]);
            }
            // the lane holding (c*, pixel) has its fire mask and maximum; the
            // lanes of the same pixel in its wave take c*'s first firing step
            // from it
            const int32_t cr = (cs % (16 * RT)) / 16, cq = (cs >> 2) & 3, ci = cs & 3;
            const bool own = cs / (16 * RT) == w && cq == q;
            uint32_t om = 0;
            int32_t ox = 0;
            float first = 0.0f;  // c*'s value at inhibit's step, from pass 1
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (cr == r && ci == i) om = mask[r][i], ox = mxb[r][i], first = gv[r][i];
            const int32_t my_t0 = om ? g.t - 32 + int32_t(__builtin_clz(om)) : 0;
            const int32_t t0 = __shfl(my_t0, cq * 16 + j, 64);
            // pass 2: the value of c* at its first firing step (convwta's t0
            // value).  Where that step is inhibit's step the value is pass 1's
            // (the same operands and sums): with cumulative spikes and
            // non-negative weights the potentials never fall, so c* fires
            // first exactly at inhibit's step and no lane needs the pass.
            if (__builtin_amdgcn_ballot_w64(own && pvalid && om != 0 && my_t0 != ti))
            {
                StepBytes<S, PAIRED> r;
                load_raw(r, pb + t0 * g.sstep, toff);
                mfma_sub(a, r, gv);
#pragma unroll
                for (int r2 = 0; r2 < RT; ++r2)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (cr == r2 && ci == i) first = gv[r2][i];
            }
            if (own && pvalid)
            {
                const uint32_t fwd = om ? __builtin_bitreverse32(om) >> (32 - g.t) : 0u;
                summary[int64_t(n) * hw + p] =
                    Summary{cs, fwd, om ? first * TWO_POW_24 : 0.0f,
                            om ? __builtin_bit_cast(float, ox) * TWO_POW_24 : 0.0f};
                smap[p] = make_uint2(uint32_t(cs), fwd);
            }
        }
        if (s_out)
        {
            if (int64_t(threadIdx.x) < s_head) s_out[threadIdx.x] = 0;
            for (int64_t i = s_head + s_units * 16 + threadIdx.x; i < s_bytes; i += blockDim.x) s_out[i] = 0;
        }
        prev = n;
    }
    if (prev >= 0 && (out || pooled))
    {
        __syncthreads();
        scatter_sample(prev, smap, out, pooled, g);
    }
}

// -- the fixed-point int8 form of the whole-sample kernel --------------------
//
// The stage's bytes feed v_mfma_i32_16x16x64_i8 with no conversion but a
// bias: byte b is staged as b ^ 0x80, which the signed int8 MFMA reads as b -
// 128 for every b in [0, 255] (the zero border holds 0x80, a true 0).  The
// weights become fixed point: w_q = rint(w * 2^e), and the fragments hold
// -w_q split into two int8 planes 256 * hi + lo (lo in [-128, 127]).  The hi
// chain starts at 0 and the lo chain at the channel's biased threshold thc =
// floor(th * 2^e) - 128 * A (A = the channel's sum of w_q), so 256 * hi_sum +
// lo_sum = thc - (v_q - 128 * A) = th_q - v_q =: d, v_q the exact integer
// potential (modulo 2^32, the accumulators' own wrap-around).  The sign bit
// of d is "v > th", and its minimum over the steps gives the maximum
// potential.  e is the
// largest exponent that keeps every |w_q| <= 32639 and th_q - v_q inside
// int32 for any input bytes (255 * the channel's positive and negative weight
// sums, with one unit of rounding slack a weight), so the test is exact for
// every U8 input.  Against FP16 weights this rounds each weight to 2^-e / 2
// (FP16: 2^-11 of the weight), and any sum of weights exact at that scale
// (the dyadic test weights) is exact.
//
// K-units: a window row (kw pixels x cin_pad channels) is contiguous in the
// stage; unit u covers 16 bytes of row u / upr at byte 16 * (u % upr), bytes
// past the row (into the next pixel) carrying zero weights.  K-step s: lane
// group q takes unit 4s + q (the i8 MFMA's lane l holds K bytes 16 * (l / 16)
// .. +16 of row or column l % 16).
typedef int32_t int4v __attribute__((ext_vector_type(4)));

constexpr int MAX_STEPS8 = 8;  // K-steps of 64 bytes -> 512

// The fixed-point scale: the largest e with max|w| * 2^e <= 32639.
__device__ __forceinline__ int32_t fixed_exponent(float m)
{
    if (!(m > 0.0f)) return 0;
    if (!(m <= 3.0e38f)) return -120;
    int x = 0;
    (void)frexpf(m, &x);  // m = f * 2^x, f in [0.5, 1)
    int32_t e = 15 - x;
    if (ldexpf(m, e) > 32639.0f) --e;
    return e < -120 ? -120 : (e > 120 ? 120 : e);
}

// pot_q > floor(th * 2^e) <=> pot_q * 2^-e > th for integers pot_q; thresholds
// past 2^30 clamp there (no potential reaches them: |pot_q| < 2^31 - 2^30 by
// the choice of e).
__device__ __forceinline__ int32_t threshold_fixed(float th, int32_t e)
{
    const float x = floorf(ldexpf(th, e));
    return x >= 1073741824.0f ? 1073741824 : int32_t(x);
}

// Lowers e until the fire test's difference th_q - v_q stays inside int32 for
// any U8 input: v_q lies in [-255 * N_q, 255 * P_q], P_q and N_q being the
// largest per-channel sums of the positive and negative w_q (each w_q within
// 1/2 of w * 2^e, so P_q <= P * 2^e + K / 2; K of slack covers that and the
// float sums), and th_q in [0, 2^30].
__device__ __forceinline__ int32_t bounded_exponent(int32_t e, float th, double pmax, double nmax, int32_t k)
{
    const double lim = 2147483647.0;
    for (; e > -120; --e)
    {
        const double s = ldexp(1.0, e);
        if (255.0 * (pmax * s + k) <= lim && double(threshold_fixed(th, e)) + 255.0 * (nmax * s + k) <= lim) break;
    }
    return e;
}

__device__ __forceinline__ int32_t unit_off(const FusedGeom &g, int32_t u)
{
    const int32_t uu = u < g.units ? u : 0;
    return (uu / g.upr) * g.pw * g.cin_pad + (uu % g.upr) * 16;
}

// The int8 header: e, then from HDR8_THC on the biased threshold thc of each
// of 256 output channels (th_q for channels past cout, whose weights are 0).
constexpr int HDR8_THC = 256;
constexpr int HDR8_BYTES = 4 * (HDR8_THC + 256);

// The int8 scale of the FP32 masters: e from max|w| and the largest
// per-channel sums of the positive and of the negative weights (hdr[0]), and
// every channel's threshold th_q (hdr[HDR8_THC + c]), from which
// prep_fragments8 subtracts 128 * A.  One block, four threads a channel,
// each reading every fourth pair of the channel's weights with all its loads
// issued at once; the sums combine in a fixed order, so e is deterministic.
constexpr int SCALE_CHUNK = 16;  // pairs of weights a thread loads at once

__global__ void __launch_bounds__(1024)
    prep_scale8(const float *__restrict__ k, int32_t *__restrict__ hdr, float th, int32_t cout, int32_t ksz)
{
    __shared__ float part[16][3];
    const int32_t sub = threadIdx.x & 3, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float m = 0.0f, pm = 0.0f, nm = 0.0f;
    const int32_t c = threadIdx.x >> 2;
    if (c < cout)
    {
        const float *kc = k + int64_t(c) * ksz;
        float ps = 0.0f, ns = 0.0f;
        if ((reinterpret_cast<uintptr_t>(kc) & 7) == 0)
            // chunks of SCALE_CHUNK pairs, each chunk's loads in flight together
            for (int32_t r0 = 0; 8 * r0 < ksz; r0 += SCALE_CHUNK)
            {
                float2 x[SCALE_CHUNK];
#pragma unroll
                for (int r = 0; r < SCALE_CHUNK; ++r)
                {
                    const int32_t i = 2 * (sub + 4 * (r0 + r));
                    x[r] = i + 1 < ksz ? *reinterpret_cast<const float2 *>(kc + i)
                                       : make_float2(i < ksz ? kc[i] : 0.0f, 0.0f);
                }
#pragma unroll
                for (int r = 0; r < SCALE_CHUNK; ++r)
                {
                    m = fmaxf(m, fmaxf(fabsf(x[r].x), fabsf(x[r].y)));
                    ps += fmaxf(x[r].x, 0.0f) + fmaxf(x[r].y, 0.0f);
                    ns += fmaxf(-x[r].x, 0.0f) + fmaxf(-x[r].y, 0.0f);
                }
            }
        else
            for (int32_t i = sub; i < ksz; i += 4)
            {
                const float x = kc[i];
                m = fmaxf(m, fabsf(x)), ps += fmaxf(x, 0.0f), ns += fmaxf(-x, 0.0f);
            }
        ps += __shfl_xor(ps, 1, 64), ns += __shfl_xor(ns, 1, 64);
        ps += __shfl_xor(ps, 2, 64), ns += __shfl_xor(ns, 2, 64);
        pm = ps, nm = ns;
    }
    m = wave_max(m), pm = wave_max(pm), nm = wave_max(nm);
    if (lane == 0) part[wv][0] = m, part[wv][1] = pm, part[wv][2] = nm;
    __syncthreads();
    __shared__ int32_t e_s;
    if (threadIdx.x == 0)
    {
        for (int i = 0; i < int(blockDim.x >> 6); ++i)
            m = fmaxf(m, part[i][0]), pm = fmaxf(pm, part[i][1]), nm = fmaxf(nm, part[i][2]);
        const int32_t e = bounded_exponent(fixed_exponent(m), th, double(pm), double(nm), ksz);
        e_s = e, hdr[0] = e;
    }
    __syncthreads();
    if (threadIdx.x < 256) hdr[HDR8_THC + threadIdx.x] = threshold_fixed(th, e_s);
}

// [row tile][K step][plane][lane] int8 fragments of the negated fixed-point
// weights -w_q = 256 * hi + lo (plane 0 = hi, 1 = lo) at the scale
// prep_scale8 chose (hdr[0]); a thread writes both
// planes of one (row tile, K step, lane), whose 16 weights all belong to one
// output channel, and subtracts 128 times their fixed-point sum from that
// channel's threshold (integer atomics: the result does not depend on their
// order), leaving thc = th_q - 128 * A.
__global__ void __launch_bounds__(64)
    prep_fragments8(const float *__restrict__ k, int4v *__restrict__ frag, int32_t *__restrict__ hdr,
                    int32_t cout, int32_t cin, int32_t kh, int32_t kw, int32_t cin_pad, int32_t upr, int32_t units,
                    int32_t steps, int32_t row_tiles)
{
    const int32_t e = hdr[0];
    const int32_t total = row_tiles * steps * 64;
    for (int32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x)
    {
        const int32_t lane = idx % 64, s = (idx / 64) % steps, rt = idx / (64 * steps);
        const int32_t o = rt * 16 + lane % 16, u = 4 * s + lane / 16, ky = u / upr;
        uint32_t hw_[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
        int32_t sum = 0;
#pragma unroll
        for (int b = 0; b < 16; ++b)
        {
            const int32_t pos = (u % upr) * 16 + b, kx = pos / cin_pad, c = pos % cin_pad;
            float x = 0.0f;
            if (o < cout && u < units && kx < kw && c < cin) x = k[((int64_t(o) * cin + c) * kh + ky) * kw + kx];
            const int32_t q = int32_t(rintf(ldexpf(x, e))), nq = -q;
            const int32_t hi = (nq + 128) >> 8, lo = nq - hi * 256;
            hw_[b / 4] |= uint32_t(hi & 0xFF) << (8 * (b % 4));
            lw[b / 4] |= uint32_t(lo & 0xFF) << (8 * (b % 4));
            sum += q;
        }
        const int64_t base = (int64_t(rt) * steps + s) * 2 * 64 + lane;
        frag[base] = int4v{int32_t(hw_[0]), int32_t(hw_[1]), int32_t(hw_[2]), int32_t(hw_[3])};
        frag[base + 64] = int4v{int32_t(lw[0]), int32_t(lw[1]), int32_t(lw[2]), int32_t(lw[3])};
        if (sum != 0) atomicSub(reinterpret_cast<uint32_t *>(hdr + HDR8_THC + o), 128u * uint32_t(sum));
    }
}

template <bool WIDE>
__device__ __forceinline__ int4v load16(const uint8_t *p)
{
    if (WIDE) return *reinterpret_cast<const int4v *>(p);
    const uint2 lo = *reinterpret_cast<const uint2 *>(p), hi = *reinterpret_cast<const uint2 *>(p + 8);
    return int4v{int32_t(lo.x), int32_t(lo.y), int32_t(hi.x), int32_t(hi.y)};
}

template <int S8, bool WIDE>
__device__ __forceinline__ void load_units(int4v (&b)[S8], const uint8_t *base, const int32_t (&off)[S8])
{
#pragma unroll
    for (int s = 0; s < S8; ++s) b[s] = load16<WIDE>(base + off[s]);
}

// One step's MFMAs: hi and lo planes of -w_q into their own accumulators,
// the lo chain started at the channels' biased thresholds thc, so 256 * ch +
// cl = thc + raw(-w_q) = th_q - v_q (see fixed_d).
template <int RT, int S8>
__device__ __forceinline__ void chain8(const int4v (&a)[RT][S8][2], const int4v (&b)[S8], const int4v (&thc)[RT],
                                       int4v (&ch)[RT], int4v (&cl)[RT])
{
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) ch[rt] = int4v{0, 0, 0, 0}, cl[rt] = thc[rt];
#pragma unroll
    for (int s = 0; s < S8; ++s)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
        {
            ch[rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[rt][s][0], b[s], ch[rt], 0, 0, 0);
            cl[rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[rt][s][1], b[s], cl[rt], 0, 0, 0);
        }
}

// th_q - v_q from a chain's accumulators (hi plane, and lo plane started at
// thc), computed modulo 2^32 like the accumulators: exact, since the choice of
// e keeps th_q - v_q inside int32.
__device__ __forceinline__ int32_t fixed_d(int32_t hi, int32_t lo)
{
    return int32_t(uint32_t(hi) * 256u + uint32_t(lo));
}

// value_update on an exact fixed-point potential: d = th_q - v_q has its sign
// bit set exactly when v > th, the running minimum of d is th_q minus the
// running maximum of v (started at v = 0), and fd keeps d of the first step
// that fired (the mask was 0 before it; later steps leave it).
__device__ __forceinline__ void value_update8(int32_t hi, int32_t lo, uint32_t &mask, int32_t &mn, int32_t &fd)
{
    const int32_t d = fixed_d(hi, lo);
    mask = __builtin_amdgcn_alignbit(mask, uint32_t(d), 31);
    mn = min(mn, d);
    fd = mask <= 1u ? d : fd;
}

// The per-value state of a tile's step loop.
template <int RT>
struct FireState8
{
    uint32_t mask[RT][4];  // bit (T-1-t): fired at step t
    int32_t mn[RT][4];     // th_q - the running maximum of the potential (from 0)
    int32_t fd[RT][4];     // th_q - the potential at the first step that fired
};

template <int RT>
__device__ __forceinline__ void update_all8(const int4v (&uh)[RT], const int4v (&ul)[RT], FireState8<RT> &f)
{
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int i = 0; i < 4; ++i) value_update8(uh[rt][i], ul[rt][i], f.mask[rt][i], f.mn[rt][i], f.fd[rt][i]);
}

// step_iter of the int8 kernel: the MFMA pairs of step k+1 on `rin`, step
// k's value updates spread between them, and the loads of step k+2's units
// into `rout` right after the MFMAs that read the same unit of `rin`.
template <int RT, int S8, bool WIDE>
__device__ __forceinline__ void step_iter8(const int4v (&a)[RT][S8][2], const int4v (&rin)[S8], int4v (&ch)[RT],
                                           int4v (&cl)[RT], const int4v (&uh)[RT], const int4v (&ul)[RT],
                                           const int4v (&thc)[RT], FireState8<RT> &f, int4v (&rout)[S8],
                                           const uint8_t *next, const int32_t (&off)[S8])
{
    constexpr int V = 4 * RT, SPAN = S8 * RT - RT;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) ch[rt] = int4v{0, 0, 0, 0}, cl[rt] = thc[rt];
#pragma unroll
    for (int s = 0; s < S8; ++s)
    {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
        {
            ch[rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[rt][s][0], rin[s], ch[rt], 0, 0, 0);
            cl[rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[rt][s][1], rin[s], cl[rt], 0, 0, 0);
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (SPAN > 0 && RT + (v * SPAN) / V == s * RT + rt)
                    value_update8(uh[v / 4][v % 4], ul[v / 4][v % 4], f.mask[v / 4][v % 4], f.mn[v / 4][v % 4],
                                  f.fd[v / 4][v % 4]);
        }
        rout[s] = load16<WIDE>(next + off[s]);
    }
    if (SPAN <= 0) update_all8(uh, ul, f);
    __builtin_amdgcn_sched_barrier(0);
}

// The whole-sample kernel in fixed-point int8 arithmetic: fused_layer's
// structure (persistent blocks over samples, staging, 16-pixel tiles, the
// steps without barriers, the two gathered tile-end passes, the in-loop zero
// stores and the scatter of 1s) with i8 MFMAs on the raw stage bytes.
template <int RT, int S8, bool WIDE, bool DIAG>
__global__ void __launch_bounds__(layer_threads<RT>())
    fused_layer8(const uint8_t *__restrict__ in, const int4v *__restrict__ frag, const int32_t *__restrict__ hdr,
                 Summary *__restrict__ summary, float *__restrict__ pot_out, uint8_t *__restrict__ fs_out,
                 uint8_t *__restrict__ out, uint8_t *__restrict__ pooled, FusedGeom g)
{
    extern __shared__ __align__(16) uint8_t lds[];
    uint8_t *stage = lds;                                                       // [t][ph][pw][cin_pad] U8
    // the waves' partials of each pixel side by side, two buffers each:
    // [2][COLS][MAX_WAVES], so a lane reads its pixel's with two 16-byte loads
    uint32_t *any_s = reinterpret_cast<uint32_t *>(lds + g.stage_bytes);
    float *red_v = reinterpret_cast<float *>(any_s + 2 * MAX_WAVES * COLS);
    int32_t *red_i = reinterpret_cast<int32_t *>(red_v + 2 * MAX_WAVES * COLS);
    uint2 *smap = reinterpret_cast<uint2 *>(red_i + 2 * MAX_WAVES * COLS);       // [ho*wo] {c*, fire mask}

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
    const int32_t hw = g.ho * g.wo;
    const int32_t o0 = w * 16 * RT + 4 * q;  // the lane's first channel; value (r, i) is channel o0 + 16r + i

    int4v a[RT][S8][2];
#pragma unroll
    for (int s = 0; s < S8; ++s)
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int p = 0; p < 2; ++p) a[r][s][p] = frag[(((w * RT + r) * S8 + s) * 2 + p) * 64 + lane];
    const int32_t e = hdr[0];
    int4v thc[RT];  // the biased thresholds of the lane's channels: the lo chains' start
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) thc[r][i] = hdr[HDR8_THC + o0 + r * 16 + i];
    __builtin_amdgcn_s_waitcnt(0);  // as in fused_layer: no pending fragment loads in the tile loop
    const float scale = ldexpf(1.0f, -e);
    const uint32_t thq = uint32_t(threshold_fixed(g.thf, e));
    int32_t off[S8];
#pragma unroll
    for (int s = 0; s < S8; ++s) off[s] = unit_off(g, 4 * s + q);
    // the border (and padded channels) hold the bias, a true 0
    constexpr uint32_t BIAS = 0x80808080u;
    for (int32_t i = threadIdx.x; i < g.stage_bytes / 16; i += blockDim.x)
        reinterpret_cast<uint4 *>(stage)[i] = make_uint4(BIAS, BIAS, BIAS, BIAS);
    // partials of waves past g.waves stay neutral: no steps, value -1
    for (int32_t i = threadIdx.x; i < 2 * MAX_WAVES * COLS; i += blockDim.x)
        any_s[i] = 0, red_v[i] = -1.0f, red_i[i] = INT32_MAX;

    int32_t prev = -1;  // the sample whose 1s are not stored yet
    for (int32_t n = blockIdx.x; n < g.n; n += gridDim.x)
    {
        __syncthreads();
        stage_sample(in, n, g, stage, BIAS);
        __syncthreads();
        if (fs_out) first_spike_map(stage, n, g, fs_out, BIAS);
        if (prev >= 0 && (out || pooled)) scatter_sample(prev, smap, out, pooled, g);

        const int64_t s_bytes = int64_t(g.t) * g.cout * hw, p_bytes = int64_t(g.t) * g.cout * g.pho * g.pwo;
        uint8_t *s_out = out ? out + int64_t(n) * s_bytes : nullptr;
        uint8_t *p_out = pooled ? pooled + int64_t(n) * p_bytes : nullptr;
        int64_t s_head = 0, s_units = 0, s_upt = 0;  // edge bytes, 16-byte units, units a tile
        if (s_out)
        {
            s_head = min(int64_t((16 - (reinterpret_cast<uintptr_t>(s_out) & 15)) & 15), s_bytes);
            s_units = (s_bytes - s_head) / 16;
            s_upt = (s_units + g.tiles - 1) / g.tiles;
        }
        uint4 *s_body = reinterpret_cast<uint4 *>(s_out + s_head);
        for (int32_t tile = 0; tile < g.tiles; ++tile)
        {
            if (p_out) zero_slice(p_out, int64_t(tile) * g.zp, min(int64_t(tile + 1) * g.zp, p_bytes));
            const int64_t u0 = int64_t(tile) * s_upt, u1 = min(u0 + s_upt, s_units);
            const int32_t buf = (tile & 1) * MAX_WAVES * COLS;
            const int32_t p = tile * COLS + j;
            const bool pvalid = p < hw;
            const int32_t pc = pvalid ? p : hw - 1, y = pc / g.wo, x = pc % g.wo;
            const uint8_t *pb = stage + (y * g.pw + x) * g.cin_pad;  // the pixel's window at step 0
            FireState8<RT> f;
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) f.mask[r][i] = 0, f.mn[r][i] = int32_t(thq), f.fd[r][i] = 0;

            {
                const int32_t last = g.t - 1;
                int4v rA[S8], rB[S8];
                int4v hA[RT], lA[RT], hB[RT], lB[RT];
                auto diag = [&](int32_t k, const int4v (&uh)[RT], const int4v (&ul)[RT]) {
                    if (DIAG && pvalid)
#pragma unroll
                        for (int r = 0; r < RT; ++r)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                            {
                                const int32_t o = o0 + r * 16 + i;
                                if (o < g.cout)
                                    pot_out[((int64_t(n) * g.t + k) * g.cout + o) * hw + p] =
                                        float(int32_t(thq - uint32_t(fixed_d(uh[r][i], ul[r][i])))) * scale;
                            }
                };
                load_units<S8, WIDE>(rA, pb, off);
                load_units<S8, WIDE>(rB, pb + min(1, last) * g.sstep, off);
                chain8(a, rA, thc, hA, lA);
                int32_t k = 0;
                int64_t zu = u0 + int32_t(threadIdx.x);
                for (; k + 2 < g.t; k += 2)
                {
                    step_iter8<RT, S8, WIDE>(a, rB, hB, lB, hA, lA, thc, f, rA, pb + (k + 2) * g.sstep, off);
                    diag(k, hA, lA);
                    if (zu < u1) zero16(s_body + zu);
                    zu += blockDim.x;
                    step_iter8<RT, S8, WIDE>(a, rA, hA, lA, hB, lB, thc, f, rB, pb + min(k + 3, last) * g.sstep,
                                             off);
                    diag(k + 1, hB, lB);
                }
                if (k + 1 < g.t)
                {
                    step_iter8<RT, S8, WIDE>(a, rB, hB, lB, hA, lA, thc, f, rA, pb + last * g.sstep, off);
                    diag(k, hA, lA);
                    update_all8(hB, lB, f);
                    diag(k + 1, hB, lB);
                }
                else
                {
                    update_all8(hA, lA, f);
                    diag(k, hA, lA);
                }
                for (; zu < u1; zu += blockDim.x) zero16(s_body + zu);
            }

            {
                uint32_t any = 0;
#pragma unroll
                for (int r = 0; r < RT; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) any |= f.mask[r][i];
                any |= __shfl_xor(any, 16, 64);
                any |= __shfl_xor(any, 32, 64);
                if (lane < COLS) any_s[buf + lane * MAX_WAVES + w] = any;
            }
            __syncthreads();
            // inhibit's kept channel c* at its step ti = T - max(cnt, 1): the
            // first maximum of the thresholded values there.  A value fired
            // at ti exactly when its mask says so.  When the pixel's fired
            // steps (all channels) form one run up to the last step, ti is the
            // first step any channel fired, so a value that fired at ti fired
            // first there and fd holds it; otherwise a pass of the step
            // loop's MFMA sequence at ti (each lane at its own pixel's step)
            // gives the values, bit for bit.  The thresholded values compare
            // as floats (exactly for |v_q| < 2^24, as with spike inputs).
            {
                const uint4 a0 = *reinterpret_cast<const uint4 *>(any_s + buf + j * MAX_WAVES);
                const uint4 a1 = *reinterpret_cast<const uint4 *>(any_s + buf + j * MAX_WAVES + 4);
                const uint32_t all = (a0.x | a0.y | a0.z | a0.w) | (a1.x | a1.y | a1.z | a1.w);
                const int32_t cnt = __builtin_popcount(all);
                const int32_t ti = g.t - (cnt > 1 ? cnt : 1), sh = g.t - 1 - ti;
                const bool run = (all & (all + 1u)) == 0u;
                int32_t dti[RT][4];
#pragma unroll
                for (int r = 0; r < RT; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) dti[r][i] = f.fd[r][i];
                if (__builtin_amdgcn_ballot_w64(!run))
                {
                    int4v r[S8], gh[RT], gl[RT];
                    load_units<S8, WIDE>(r, pb + ti * g.sstep, off);
                    chain8(a, r, thc, gh, gl);
#pragma unroll
                    for (int r2 = 0; r2 < RT; ++r2)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (!run) dti[r2][i] = fixed_d(gh[r2][i], gl[r2][i]);
                }
                float bv = -1.0f;
                int32_t bi = INT32_MAX;
#pragma unroll
                for (int r2 = 0; r2 < RT; ++r2)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                    {
                        const float xp = (f.mask[r2][i] >> sh) & 1u ? float(int32_t(thq - uint32_t(dti[r2][i]))) : 0.0f;
                        const bool take = xp > bv;
                        bv = take ? xp : bv;
                        bi = take ? o0 + r2 * 16 + i : bi;
                    }
                combine(bv, bi, __shfl_xor(bv, 16, 64), __shfl_xor(bi, 16, 64));
                combine(bv, bi, __shfl_xor(bv, 32, 64), __shfl_xor(bi, 32, 64));
                if (lane < COLS) red_v[buf + lane * MAX_WAVES + w] = bv, red_i[buf + lane * MAX_WAVES + w] = bi;
            }
            __syncthreads();
            int32_t cs;
            {
                // the waves' partials of the pixel, combined as a tree (the
                // order (value desc, channel asc) is total)
                const float4 v0 = *reinterpret_cast<const float4 *>(red_v + buf + j * MAX_WAVES);
                const float4 v1 = *reinterpret_cast<const float4 *>(red_v + buf + j * MAX_WAVES + 4);
                const int4 i0 = *reinterpret_cast<const int4 *>(red_i + buf + j * MAX_WAVES);
                const int4 i1 = *reinterpret_cast<const int4 *>(red_i + buf + j * MAX_WAVES + 4);
                float m0 = v0.x, m1 = v0.z, m2 = v1.x, m3 = v1.z;
                int32_t c0 = i0.x, c1 = i0.z, c2 = i1.x, c3 = i1.z;
                combine(m0, c0, v0.y, i0.y), combine(m1, c1, v0.w, i0.w);
                combine(m2, c2, v1.y, i1.y), combine(m3, c3, v1.w, i1.w);
                combine(m0, c0, m1, c1), combine(m2, c2, m3, c3);
                combine(m0, c0, m2, c2);
                cs = c0;
            }
            const int32_t cr = (cs % (16 * RT)) / 16, cq = (cs >> 2) & 3, ci = cs & 3;
            const bool own = cs / (16 * RT) == w && cq == q;
            uint32_t om = 0;
            int32_t ox = 0, first = 0;  // th_q - c*'s maximum and its value at its first firing step
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (cr == r && ci == i) om = f.mask[r][i], ox = f.mn[r][i], first = f.fd[r][i];
            if (own && pvalid)
            {
                const uint32_t fwd = om ? __builtin_bitreverse32(om) >> (32 - g.t) : 0u;
                summary[int64_t(n) * hw + p] =
                    Summary{cs, fwd, om ? float(int32_t(thq - uint32_t(first))) * scale : 0.0f,
                            om ? float(int32_t(thq - uint32_t(ox))) * scale : 0.0f};
                smap[p] = make_uint2(uint32_t(cs), fwd);
            }
        }
        if (s_out)
        {
            if (int64_t(threadIdx.x) < s_head) s_out[threadIdx.x] = 0;
            for (int64_t i = s_head + s_units * 16 + threadIdx.x; i < s_bytes; i += blockDim.x) s_out[i] = 0;
        }
        prev = n;
    }
    if (prev >= 0 && (out || pooled))
    {
        __syncthreads();
        scatter_sample(prev, smap, out, pooled, g);
    }
}

// convwta over the summaries: candidates are the kept (channel, pixel) pairs;
// value (T-1-t0) + first/diff with min 0 and diff = 2*max (wta.cpp:49-64);
// k rounds of first maximum in flat (c, y, x) order with box and channel
// masking (wta.cpp:128-150).  One block per sample; candidates in LDS.
// wkey[slot] = channel << 16 | pixel of winner `slot` (sample * count +
// rank), -1 past the sample's winners: the compact list the STDP reads.
__global__ void __launch_bounds__(256) fused_select(const Summary *__restrict__ summary, int32_t t, int32_t ho,
                                                    int32_t wo, int32_t ry, int32_t rx, int32_t count,
                                                    spk_winner *__restrict__ win, int64_t *__restrict__ counts,
                                                    int32_t *__restrict__ wkey)
{
    extern __shared__ __align__(16) uint8_t lds[];
    const int32_t hw = ho * wo;
    float *val = reinterpret_cast<float *>(lds);
    int32_t *chan = reinterpret_cast<int32_t *>(val + hw);
    __shared__ float mm[32];
    __shared__ ArgMax am[16];
    const int64_t s = blockIdx.x;
    const Summary *sm = summary + s * hw;
    // per-sample maximum of the inhibited potentials
    float m = 0.0f;
    for (int32_t p = threadIdx.x; p < hw; p += blockDim.x) m = m < sm[p].max ? sm[p].max : m;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) mm[threadIdx.x >> 6] = m;
    __syncthreads();
    m = mm[0];
    for (int i = 1; i < int(blockDim.x >> 6); ++i) m = m < mm[i] ? mm[i] : m;
    const float diff = (m - 0.0f) * 2.0f;
    for (int32_t p = threadIdx.x; p < hw; p += blockDim.x)
    {
        const Summary e = sm[p];
        chan[p] = e.c;
        val[p] = e.mask ? float(t - int32_t(__builtin_ctz(e.mask)) - 1) + (e.first - 0.0f) / diff : -FLT_MAX;
    }
    __syncthreads();
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMax a{-INFINITY, INT32_MAX};
        for (int32_t p = threadIdx.x; p < hw; p += blockDim.x)
            if (better(val[p], chan[p] * hw + p, a.v, a.i)) a.v = val[p], a.i = chan[p] * hw + p;
        a = block_argmax(a, am);
        if (fabsf(a.v - (-FLT_MAX)) < 1e-5f) break;
        const int32_t z = a.i / hw, y = (a.i % hw) / wo, x = (a.i % hw) % wo;
        if (threadIdx.x == 0)
            win[s * count + k] = spk_winner{0, 0, z, y, x}, wkey[s * count + k] = z << 16 | (a.i % hw);
        for (int32_t p = threadIdx.x; p < hw; p += blockDim.x)
        {
            const int32_t py = p / wo, px = p % wo;
            if (chan[p] == z || (abs(py - y) <= ry && abs(px - x) <= rx)) val[p] = -FLT_MAX;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[s] = k;
    for (int32_t r = k + int32_t(threadIdx.x); r < count; r += blockDim.x)
        wkey[s * count + r] = -1, win[s * count + r] = spk_winner{0, 0, 0, 0, 0};
}

// The first maximum of (v, then smallest i) over the wave, in every lane:
// butterflies over quads (quad_perm), eights (row_half_mirror) and rows
// (row_mirror), then rows 0..3 chained by row_bcast15/row_bcast31 into lane
// 63, read back as a uniform value.  All 64 lanes must be active.
__device__ __forceinline__ void wave_argmax_dpp(float &v, int32_t &i)
{
    auto step = [&](auto ctrl, auto rows) {
        const int32_t vb = __builtin_bit_cast(int32_t, v);
        const float ov = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(vb, vb, decltype(ctrl)::value,
                                                                               decltype(rows)::value, 0xF, false));
        const int32_t oi = __builtin_amdgcn_update_dpp(i, i, decltype(ctrl)::value, decltype(rows)::value, 0xF, false);
        if (better(ov, oi, v, i)) v = ov, i = oi;
    };
    step(std::integral_constant<int, 0xB1>{}, std::integral_constant<int, 0xF>{});   // lane ^ 1
    step(std::integral_constant<int, 0x4E>{}, std::integral_constant<int, 0xF>{});   // lane ^ 2
    step(std::integral_constant<int, 0x141>{}, std::integral_constant<int, 0xF>{});  // 7 - lane within 8
    step(std::integral_constant<int, 0x140>{}, std::integral_constant<int, 0xF>{});  // 15 - lane within 16
    step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{});  // rows 1, 3 take 15, 47
    step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{});  // rows 2, 3 take 31
    v = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), 63));
    i = __builtin_amdgcn_readlane(i, 63);
}

// fused_select with one wave per sample, for planes of at most 64 * VPL
// pixels: pixel p = lane + 64 i is the lane's value i, each round's first
// maximum is a wave reduction with no barrier, and the unused winner slots
// are zeroed.  Same arithmetic and tie order as fused_select.
template <int VPL>
__global__ void __launch_bounds__(256) fused_select_wave(const Summary *__restrict__ summary, int32_t t, int32_t ho,
                                                         int32_t wo, int32_t ry, int32_t rx, int32_t count,
                                                         spk_winner *__restrict__ win, int64_t *__restrict__ counts,
                                                         int32_t *__restrict__ wkey, int64_t n)
{
    const int64_t s = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (s >= n) return;
    const int lane = threadIdx.x & 63;
    const int32_t hw = ho * wo;
    const Summary *sm = summary + s * hw;
    Summary e[VPL];
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < VPL; ++i)
    {
        const int32_t p = lane + 64 * i;
        e[i] = p < hw ? sm[p] : Summary{-1, 0, 0.0f, 0.0f};
        m = m < e[i].max ? e[i].max : m;
    }
    m = wave_max(m);
    const float diff = (m - 0.0f) * 2.0f;
    float val[VPL];
    int32_t chan[VPL], py[VPL], px[VPL];
#pragma unroll
    for (int i = 0; i < VPL; ++i)
    {
        const int32_t p = lane + 64 * i;
        chan[i] = e[i].c;
        val[i] = p >= hw ? -INFINITY
                         : e[i].mask ? float(t - int32_t(__builtin_ctz(e[i].mask)) - 1) + (e[i].first - 0.0f) / diff
                                     : -FLT_MAX;
        // slots past the plane sit far outside every box (and hold channel -1)
        py[i] = p < hw ? p / wo : -(1 << 24), px[i] = p < hw ? p % wo : 0;
    }
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMax a{-INFINITY, INT32_MAX};
#pragma unroll
        for (int i = 0; i < VPL; ++i)
            if (better(val[i], chan[i] * hw + lane + 64 * i, a.v, a.i)) a.v = val[i], a.i = chan[i] * hw + lane + 64 * i;
        wave_argmax_dpp(a.v, a.i);
        if (fabsf(a.v - (-FLT_MAX)) < 1e-5f) break;
        const int32_t z = a.i / hw, y = (a.i % hw) / wo, x = (a.i % hw) % wo;
        if (lane == 0) win[s * count + k] = spk_winner{0, 0, z, y, x}, wkey[s * count + k] = z << 16 | (a.i % hw);
        // a select on bitwise conditions: the short-circuit form of this test
        // was compiled with the box condition dropped (only the channel test
        // masked) by this toolchain
#pragma unroll
        for (int i = 0; i < VPL; ++i)
        {
            const bool drop = (chan[i] == z) | ((abs(py[i] - y) <= ry) & (abs(px[i] - x) <= rx));
            val[i] = drop ? -FLT_MAX : val[i];
        }
    }
    if (lane == 0) counts[s] = k;
    for (int32_t r = k + lane; r < count; r += 64)
        wkey[s * count + r] = -1, win[s * count + r] = spk_winner{0, 0, 0, 0, 0};
}

}  // namespace

KernelTimer &timer()
{
    static KernelTimer *t = new KernelTimer;
    return *t;
}

SideStream::SideStream(hipStream_t main_, bool use) : main(main_), stream(main_)
{
    if (!use) return;
    // one non-blocking side stream and a fork and a join event per device,
    // at the default priority; an event is recorded again by every call (a
    // wait takes the record that precedes it)
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    static std::vector<hipEvent_t> forks, joins;
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lock(mu);
        if (int(streams.size()) <= dev)
        {
            streams.resize(size_t(dev) + 1, nullptr);
            forks.resize(size_t(dev) + 1, nullptr);
            joins.resize(size_t(dev) + 1, nullptr);
        }
        if (!streams[size_t(dev)])
        {
            SPK_HIP(hipStreamCreateWithFlags(&streams[size_t(dev)], hipStreamNonBlocking));
            SPK_HIP(hipEventCreateWithFlags(&forks[size_t(dev)], hipEventDisableTiming));
            SPK_HIP(hipEventCreateWithFlags(&joins[size_t(dev)], hipEventDisableTiming));
        }
        stream = streams[size_t(dev)], fork = forks[size_t(dev)], done = joins[size_t(dev)];
        SPK_HIP(hipEventRecord(fork, main));
        SPK_HIP(hipStreamWaitEvent(stream, fork, 0));
    }
}

void SideStream::again()
{
    if (stream == main) return;
    SPK_HIP(hipEventRecord(fork, main));
    SPK_HIP(hipStreamWaitEvent(stream, fork, 0));
}

void SideStream::join()
{
    if (stream == main || joined) return;
    SPK_HIP(hipEventRecord(done, stream));
    SPK_HIP(hipStreamWaitEvent(main, done, 0));
    joined = true;
}

SideStream::~SideStream()
{
    // an error path: still order the call's stream after the side work
    if (stream != main && !joined && hipEventRecord(done, stream) == hipSuccess) (void)hipStreamWaitEvent(main, done, 0);
}

namespace
{
template <int RT, int S, bool PAIRED, bool DIAG>
void launch_layer(const uint8_t *in, const half8 *frag, Summary *summary, float *pot, uint8_t *fs, uint8_t *out,
                  uint8_t *pooled, const FusedGeom &g, size_t lds, hipStream_t s)
{
    static bool configured = false;  // per instantiation; the attribute is per function
    if (!configured)
    {
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_layer<RT, S, PAIRED, DIAG>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
        configured = true;
    }
    const int cus = cu_count();
    const int grid = g.n < cus ? g.n : cus;
    KernelTimer &tm = timer();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tm.on)
    {
        SPK_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
        SPK_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        SPK_HIP(hipEventRecord(e0, s));
    }
    fused_layer<RT, S, PAIRED, DIAG><<<dim3(grid), dim3(64 * g.waves), lds, s>>>(in, frag, summary, pot, fs, out,
                                                                                pooled, g);
    SPK_LAUNCHED();
    if (tm.on)
    {
        SPK_HIP(hipEventRecord(e1, s));
        tm.events.emplace_back(e0, e1);
    }
}

// The whole-sample kernel's arithmetic (spk_set_fused_arith): 0 = int8
// fixed point where its K-units fit (else FP16), 1 = FP16 operands.
std::atomic<int32_t> g_arith{0};

template <int RT, int S8, bool WIDE, bool DIAG>
void launch_layer8(const uint8_t *in, const int4v *frag, const int32_t *hdr, Summary *summary, float *pot,
                   uint8_t *fs, uint8_t *out, uint8_t *pooled, const FusedGeom &g, size_t lds, hipStream_t s)
{
    static bool configured = false;
    if (!configured)
    {
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_layer8<RT, S8, WIDE, DIAG>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
        configured = true;
    }
    const int cus = cu_count();
    const int grid = g.n < cus ? g.n : cus;
    KernelTimer &tm = timer();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tm.on)
    {
        SPK_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
        SPK_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        SPK_HIP(hipEventRecord(e0, s));
    }
    fused_layer8<RT, S8, WIDE, DIAG><<<dim3(grid), dim3(64 * g.waves), lds, s>>>(in, frag, hdr, summary, pot, fs, out,
                                                                              pooled, g);
    SPK_LAUNCHED();
    if (tm.on)
    {
        SPK_HIP(hipEventRecord(e1, s));
        tm.events.emplace_back(e0, e1);
    }
}

// DIAG (the potentials output) is built for 32 channels a wave only
template <int RT>
void dispatch_layer8(int steps, bool wide, const uint8_t *in, const int4v *frag, const int32_t *hdr,
                     Summary *summary, float *pot, uint8_t *fs, uint8_t *out, uint8_t *pooled, const FusedGeom &g,
                     size_t lds, hipStream_t s)
{
    switch (steps)
    {
#define CASE(k)                                                                                                   \
    case k:                                                                                                       \
        if (wide)                                                                                                 \
            return pot ? launch_layer8<RT, k, true, RT == 2>(in, frag, hdr, summary, pot, fs, out, pooled, g, lds, s) \
                       : launch_layer8<RT, k, true, false>(in, frag, hdr, summary, pot, fs, out, pooled, g, lds, s); \
        return pot ? launch_layer8<RT, k, false, RT == 2>(in, frag, hdr, summary, pot, fs, out, pooled, g, lds, s)    \
                   : launch_layer8<RT, k, false, false>(in, frag, hdr, summary, pot, fs, out, pooled, g, lds, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: SPK_ARG(false, "Fused layer: int8 K-steps out of range.");
    }
}

template <int RT>
void dispatch_layer(int steps, bool paired, const uint8_t *in, const half8 *frag, Summary *summary, float *pot,
                    uint8_t *fs, uint8_t *out, uint8_t *pooled, const FusedGeom &g, size_t lds, hipStream_t s)
{
    switch (steps)
    {
#define CASE(k) \
    case k:                                                                                        \
        if (paired)                                                                                \
            return pot ? launch_layer<RT, k, true, RT == 2>(in, frag, summary, pot, fs, out, pooled, g, lds, s)   \
                       : launch_layer<RT, k, true, false>(in, frag, summary, pot, fs, out, pooled, g, lds, s);    \
        return pot ? launch_layer<RT, k, false, RT == 2>(in, frag, summary, pot, fs, out, pooled, g, lds, s)      \
                   : launch_layer<RT, k, false, false>(in, frag, summary, pot, fs, out, pooled, g, lds, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: SPK_ARG(false, "Fused layer: K-steps out of range.");
    }
}
// The whole-sample kernel's channels a wave: spk_set_fused_channels forces
// 32 or 64 (0: the tuned choice, else 32); spk_set_fused_tuning(1) (the
// default) times both on the first call of each shape and keeps the faster.
std::atomic<int32_t> g_tune{1};
std::atomic<int32_t> g_force_channels{0};

using TuneKey = std::array<int64_t, 12>;
struct Tuned
{
    std::mutex mu;
    std::map<TuneKey, int> rt;
};
Tuned &tuned()
{
    static Tuned t;
    return t;
}

TuneKey tune_key(bool use8, const FusedGeom &g)
{
    return TuneKey{use8 ? 8 : 16, g.t, g.cin, g.h, g.w, g.cout, g.kh, g.kw, g.pt, g.pl, g.ph, g.pw};
}

// The channels a wave (as rt = channels / 16) for this call.  Tuning runs
// both choices on the first min(n, 2 * CUs) samples (outputs the call
// rewrites afterwards), twice each after a warm run, timed with events;
// the two produce the same bytes (each potential is the same chain of
// MFMAs), so the choice is about time only.  No tuning under a stream
// capture, for the allocation-free entry point or with the potentials output
// (built for 32 channels a wave only).
template <typename SetRt, typename Run>
int choose_channels(bool use8, const FusedGeom &g, bool potentials, bool noalloc, hipStream_t s, SetRt &&set_rt,
                    Run &&run)
{
    if (potentials) return 2;
    const int forced = g_force_channels.load();
    if (forced) return forced / 16;
    const TuneKey key = tune_key(use8, g);
    Tuned &t = tuned();
    {
        std::lock_guard<std::mutex> lock(t.mu);
        auto it = t.rt.find(key);
        if (it != t.rt.end()) return it->second;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SPK_HIP(hipStreamIsCapturing(s, &cap));
    if (!g_tune.load() || noalloc || cap != hipStreamCaptureStatusNone) return 2;
    FusedGeom gt = g;
    gt.n = std::min<int32_t>(g.n, 2 * cu_count());
    hipEvent_t e0, e1;
    SPK_HIP(hipEventCreate(&e0));
    SPK_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    int choice = 2;
    for (int r : {2, 4})
    {
        set_rt(r);
        FusedGeom gr = gt;
        gr.waves = g.waves, gr.stage_fast = g.stage_fast;
        run(r, gr);  // warm
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep)
        {
            SPK_HIP(hipEventRecord(e0, s));
            run(r, gr);
            SPK_HIP(hipEventRecord(e1, s));
            SPK_HIP(hipEventSynchronize(e1));
            float m = 0.0f;
            SPK_HIP(hipEventElapsedTime(&m, e0, e1));
            ms = rep == 0 ? m : std::min(ms, m);
        }
        if (r == 2 || ms < best) best = ms, choice = r;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lock(t.mu);
    t.rt[key] = choice;
    return choice;
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_set_fused_arith(int32_t mode)
{
    return guard([&] {
        SPK_ARG(mode == 0 || mode == 1, "Fused layer: arithmetic mode must be 0 (int8 fixed point) or 1 (FP16).");
        g_arith.store(mode);
    });
}

int32_t spk_set_fused_tuning(int32_t mode)
{
    return guard([&] {
        SPK_ARG(mode == 0 || mode == 1, "Fused layer: tuning mode must be 0 (off) or 1 (time on first call).");
        g_tune.store(mode);
    });
}

int32_t spk_set_fused_channels(int32_t channels)
{
    return guard([&] {
        SPK_ARG(channels == 0 || channels == 32 || channels == 64,
                "Fused layer: channels a wave must be 0 (tuned), 32 or 64.");
        g_force_channels.store(channels);
    });
}

int32_t spk_fused_tuned_channels(int64_t t, int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw,
                                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, int32_t *channels)
{
    return guard([&] {
        SPK_ARG(channels, "Fused layer: channels must not be null.");
        FusedGeom g{};
        g.t = int32_t(t), g.cin = int32_t(cin), g.h = int32_t(h), g.w = int32_t(w), g.cout = int32_t(cout);
        g.kh = int32_t(kh), g.kw = int32_t(kw), g.pt = int32_t(pt), g.pl = int32_t(pl);
        g.ph = int32_t(h + pt + pb), g.pw = int32_t(w + pl + pr);
        const int64_t cin_pad = cdiv(cin, 8) * 8, upr = cdiv(kw * cin_pad, 16);
        const bool use8 = g_arith.load() == 0 && cdiv(kh * upr, 4) <= MAX_STEPS8;
        Tuned &tt = tuned();
        std::lock_guard<std::mutex> lock(tt.mu);
        auto it = tt.rt.find(tune_key(use8, g));
        *channels = it == tt.rt.end() ? 0 : 16 * it->second;
    });
}

int32_t spk_kernel_timer(int32_t enable)
{
    return guard([&] {
        KernelTimer &tm = timer();
        for (auto &e : tm.events) (void)hipEventDestroy(e.first), (void)hipEventDestroy(e.second);
        tm.events.clear();
        tm.on = enable != 0;
    });
}

int32_t spk_kernel_timer_read(double *total_ms, int64_t *launches)
{
    return guard([&] {
        KernelTimer &tm = timer();
        double sum = 0;
        for (auto &e : tm.events)
        {
            SPK_HIP(hipEventSynchronize(e.second));
            float ms = 0;
            SPK_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            sum += ms;
        }
        *total_ms = sum;
        *launches = int64_t(tm.events.size());
    });
}

int32_t spk_stdp_layer_fused(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, void *kernel,
                             int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                             double threshold, int64_t ry, int64_t rx, int64_t count,
                             const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled, int64_t pkh,
                             int64_t pkw, int64_t psh, int64_t psw, spk_winner *winners, int64_t *counts,
                             void *potentials, void *stream)
{
    return guard([&] {
        stdp_layer_fused(in, n, t, cin, h, w, kernel, cout, kh, kw, pt, pl, pb, pr, threshold, ry, rx, count, configs,
                         nconfigs, out, pooled, pkh, pkw, psh, psw, winners, counts, potentials, stream, nullptr, 0);
    });
}
}

namespace spk
{
void stdp_layer_fused(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, void *kernel,
                      int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                      double threshold, int64_t ry, int64_t rx, int64_t count, const spk_stdp_config *configs,
                      int64_t nconfigs, void *out, void *pooled, int64_t pkh, int64_t pkw, int64_t psh, int64_t psw,
                      spk_winner *winners, int64_t *counts, void *potentials, void *stream, uint8_t *scratch,
                      int64_t scratch_bytes)
{
    {
        SPK_ARG(n >= 1 && t >= 1 && cin >= 1 && h >= 1 && w >= 1 && cout >= 2 && kh >= 1 && kw >= 1,
                "Fused layer: invalid shapes.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        SPK_UNSUP(t <= 32, "Fused layer: at most 32 time steps.");
        SPK_UNSUP(cout <= 32 * MAX_WAVES, "Fused layer: at most " << 32 * MAX_WAVES << " output channels.");
        SPK_UNSUP(threshold >= 0.0, "Fused layer: the conv threshold must be >= 0.");
        SPK_ARG(ry >= 0 && rx >= 0 && count >= 1, "Fused layer: invalid convwta parameters.");
        SPK_ARG(winners && counts, "Fused layer: winner buffers are required.");
        const int64_t ho = h + pt + pb - kh + 1, wo = w + pl + pr - kw + 1;
        SPK_ARG(ho >= 1 && wo >= 1, "Input and kernel shapes are not compatible.");
        SPK_ARG(n < (1LL << 31), "Fused layer: too many samples.");
        const bool learn = nconfigs > 0;
        int64_t pho = 0, pwo = 0;
        if (pooled)
        {
            SPK_ARG(pkh >= 1 && pkw >= 1 && psh >= 1 && psw >= 1, "Pool: kernel and stride must be positive.");
            pho = (ho - pkh) / psh + 1, pwo = (wo - pkw) / psw + 1;
            SPK_ARG(ho >= pkh && wo >= pkw, "Pool: input and kernel shapes are not compatible.");
        }

        FusedGeom g{};
        g.n = int32_t(n), g.t = int32_t(t), g.cin = int32_t(cin), g.h = int32_t(h), g.w = int32_t(w);
        g.cout = int32_t(cout), g.kh = int32_t(kh), g.kw = int32_t(kw), g.pt = int32_t(pt), g.pl = int32_t(pl);
        g.ho = int32_t(ho), g.wo = int32_t(wo);
        g.ph = int32_t(h + pt + pb), g.pw = int32_t(w + pl + pr);
        g.cin_pad = int32_t(cdiv(cin, 8) * 8);
        g.groups = int32_t(kh * kw * (g.cin_pad / 8));
        g.steps = int32_t(cdiv(g.groups, 4));
        g.tiles = int32_t(cdiv(ho * wo, COLS));
        // the zero-bordered stage holds every tap of every output pixel
        const int64_t sstep = int64_t(g.ph) * g.pw * g.cin_pad;
        const int64_t stage_bytes = cdiv(int64_t(t) * sstep, 16) * 16;
        // potentials come out of the MFMAs times 2^-24 (spike bytes enter as
        // FP16 subnormals); the threshold is scaled to match (exact)
        g.th = float(threshold) * (1.0f / TWO_POW_24);
        g.thf = float(threshold);
        // int8 K-units: 16 bytes of a window row (kw pixels x cin_pad channels)
        g.upr = int32_t(cdiv(int64_t(kw) * g.cin_pad, 16)), g.units = int32_t(kh) * g.upr;
        const int32_t steps8 = int32_t(cdiv(g.units, 4));
        const bool use8 = g_arith.load() == 0 && steps8 <= MAX_STEPS8;
        if (!use8)
        {
            SPK_UNSUP(threshold_scales(threshold), "Fused layer: the threshold is too small to scale exactly.");
            SPK_UNSUP(g.steps <= MAX_STEPS, "Fused layer: Cin*Kh*Kw too large (K <= 512 after padding).");
        }
        // stage + any/argmax partials (two buffers each) + the per-pixel map of kept channels and fire masks
        const int64_t lds = stage_bytes + 3 * 2 * MAX_WAVES * COLS * 4 + ho * wo * 8;
        if (lds > 163840 || ho * wo > 8192)
        {
            SPK_ARG(!scratch, "Fused layer: a sample that does not fit the LDS whole takes the region path, whose "
                              "plan allocates its buffers; use the allocating entry point for it.");
            // the sample does not fit the LDS whole: the region path
            const StdpConfigs rcfg = learn ? pack_configs(configs, nconfigs) : StdpConfigs{};
            region_layer_step((const uint8_t *)in, n, t, cin, h, w, (float *)kernel, cout, kh, kw, pt, pl, pb, pr,
                              threshold, ry, rx, count, learn ? &rcfg.c[0] : nullptr, (uint8_t *)out,
                              (uint8_t *)pooled, pkh, pkw, psh, psw, winners, counts, (float *)potentials, S(stream));
            return;
        }
        g.sstep = int32_t(sstep), g.stage_bytes = int32_t(stage_bytes);
        g.pkh = pooled ? int32_t(pkh) : 1, g.pkw = pooled ? int32_t(pkw) : 1;
        g.psh = pooled ? int32_t(psh) : 1, g.psw = pooled ? int32_t(psw) : 1;
        g.pho = int32_t(pho), g.pwo = int32_t(pwo);
        // each tile zeroes 1/tiles of the sample's pooled block, in 16-byte multiples
        SPK_ARG(t * cout * ho * wo < (1LL << 31), "Fused layer: a sample's spikes exceed 2 GiB.");
        g.zp = int32_t(cdiv(cdiv(t * cout * pho * pwo, g.tiles), 16) * 16);
        // 16 * rt channels a wave: rt 2 (32 channels, two waves per SIMD, the
        // default) or rt 4 (64 channels, one wave per SIMD); the tuner times
        // both on the first call of a shape
        auto set_rt = [&](int r) {
            g.waves = int32_t(cdiv(cout, 16 * r));
            g.stage_fast = (h * w) % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 3) == 0 &&
                           (16 * g.waves) % (g.cin_pad / 4) == 0 && 16 * g.waves / (g.cin_pad / 4) <= h * w / 4;
        };
        set_rt(2);
        SPK_ARG(int64_t(cout) * ho * wo < (1LL << 31), "Fused layer: output plane too large.");

        const StdpConfigs cfg = learn ? pack_configs(configs, nconfigs) : StdpConfigs{};
        SPK_ARG(!learn || n * count < (1LL << 31), "Fused layer: too many winner slots.");

        hipStream_t s = S(stream);
        const int64_t hw = ho * wo, hwi = h * w;
        // fragment rows for either choice of channels a wave
        const int64_t row_tiles = std::max<int64_t>(cdiv(cout, 32) * 2, cdiv(cout, 64) * 4), ksz = cin * kh * kw;
        const FusedStdp fp{n * count, int32_t(count), g.cin, g.cin_pad, g.h, g.w, g.kh, g.kw, g.pt, g.pl,
                           g.wo, int32_t(hw), g.t, int32_t(cdiv(ksz, 32))};
        // the call's buffers, carved from the stream's workspace: weight
        // fragments (FP16, or the int8 header and planes), summaries,
        // winner keys, first-spike map, STDP bit records
        const int64_t frag_bytes = use8 ? HDR8_BYTES + row_tiles * steps8 * 2 * 64 * 16 : row_tiles * g.steps * 64 * 16;
        const int64_t sizes[5] = {frag_bytes, n * hw * int64_t(sizeof(Summary)),
                                  n * count * int64_t(sizeof(int32_t)), learn ? n * hwi * g.cin_pad : 0,
                                  learn ? n * count * fp.words * int64_t(sizeof(uint32_t)) : 0};
        int64_t offs[6] = {0};
        for (int i = 0; i < 5; ++i) offs[i + 1] = offs[i] + cdiv(sizes[i], 256) * 256;
        SPK_ARG(!scratch || offs[5] <= scratch_bytes,
                "Fused layer: this call needs " << offs[5] << " bytes of scratch, more than the " << scratch_bytes
                                                << " available; split the batch.");
        uint8_t *ws = scratch ? scratch : static_cast<uint8_t *>(stream_workspace(0, s, offs[5]));
        half8 *frag = reinterpret_cast<half8 *>(ws + offs[0]);
        Summary *summary = reinterpret_cast<Summary *>(ws + offs[1]);
        int32_t *wkey = reinterpret_cast<int32_t *>(ws + offs[2]);
        uint8_t *fsm = learn ? ws + offs[3] : nullptr;
        uint32_t *bits = reinterpret_cast<uint32_t *>(ws + offs[4]);
        // prepares the weight fragments for rt and runs the whole-sample kernel
        // over the first gg.n samples
        int32_t *hdr = reinterpret_cast<int32_t *>(ws + offs[0]);
        int4v *frag8 = reinterpret_cast<int4v *>(ws + offs[0] + HDR8_BYTES);
        auto run_layer = [&](int r, const FusedGeom &gg) {
            const int64_t rtiles = int64_t(gg.waves) * r;
            if (use8)
            {
                // the header (scale exponent, biased thresholds), the fragments after
                const int64_t entries = rtiles * steps8 * 64;
                prep_scale8<<<1, 1024, 0, s>>>((const float *)kernel, hdr, g.thf, g.cout, int32_t(ksz));
                SPK_LAUNCHED();
                prep_fragments8<<<dim3(unsigned(cdiv(entries, 64))), 64, 0, s>>>(
                    (const float *)kernel, frag8, hdr, g.cout, g.cin, g.kh, g.kw, g.cin_pad, g.upr, g.units, steps8,
                    int32_t(rtiles));
                SPK_LAUNCHED();
                // the kernel also writes the spikes and pooled spikes of each sample
                (r == 4 ? dispatch_layer8<4> : dispatch_layer8<2>)(
                    steps8, g.cin_pad % 16 == 0, (const uint8_t *)in, frag8, hdr, summary, (float *)potentials, fsm,
                    (uint8_t *)out, (uint8_t *)pooled, gg, size_t(lds), s);
            }
            else
            {
                // pairs of K-steps share 16-byte reads when 16-byte units stay inside a tap
                const bool paired = g.cin_pad % 16 == 0 && g.steps >= 2;
                prep_fragments<<<dim3(unsigned(cdiv(rtiles * g.steps * 64, 256))), 256, 0, s>>>(
                    (const float *)kernel, frag, g.cout, g.cin, g.kh, g.kw, g.cin_pad, g.groups, g.steps,
                    int32_t(rtiles), paired ? 1 : 0);
                SPK_LAUNCHED();
                (r == 4 ? dispatch_layer<4> : dispatch_layer<2>)(g.steps, paired, (const uint8_t *)in, frag, summary,
                                                                  (float *)potentials, fsm, (uint8_t *)out,
                                                                  (uint8_t *)pooled, gg, size_t(lds), s);
            }
        };
        const int rt = choose_channels(use8, g, potentials != nullptr, scratch != nullptr, s, set_rt, run_layer);
        set_rt(rt);
        run_layer(rt, g);
        // convwta: one wave per sample when the plane fits a wave's registers
        if (hw <= 256)
            fused_select_wave<4><<<dim3(unsigned(cdiv(n, 4))), 256, 0, s>>>(summary, g.t, g.ho, g.wo, int32_t(ry),
                                                                             int32_t(rx), int32_t(count), winners,
                                                                             counts, wkey, n);
        else if (hw <= 1024)
            fused_select_wave<16><<<dim3(unsigned(cdiv(n, 4))), 256, 0, s>>>(summary, g.t, g.ho, g.wo, int32_t(ry),
                                                                              int32_t(rx), int32_t(count), winners,
                                                                              counts, wkey, n);
        else
            fused_select<<<dim3(unsigned(n)), 256, size_t(hw) * 8, s>>>(summary, g.t, g.ho, g.wo, int32_t(ry),
                                                                        int32_t(rx), int32_t(count), winners, counts,
                                                                        wkey);
        SPK_LAUNCHED();
        if (learn)
        {
            fused_records<<<dim3(unsigned(cdiv(n * count, 4))), 256, 0, s>>>(fsm, summary, wkey, fp, bits);
            SPK_LAUNCHED();
            launch_fused_stdp(dim3(unsigned(cout), unsigned(cdiv(ksz, STDP_THREADS))), s, (float *)kernel, bits, wkey,
                              fp, cfg.c[0]);
            SPK_LAUNCHED();
        }
    }
}
}  // namespace spk
