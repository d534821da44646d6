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
//     [t][y][x][cin_pad] (channel innermost);
//   - columns are 16 row-major output pixels.  For a column tile, the FP16
//     windows of a chunk of steps are widened once for the block into
//     [step][row][cin_pad/8 planes of padded width], row pitch = wo mod 16 so
//     the 16 pixels of one B fragment fall in distinct LDS banks; every
//     wave then runs the chunk's steps without barriers, loading the next
//     step's B fragments while it works on the current one;
//   - per step and (channel, pixel) a lane only shifts "v > th" into a mask
//     and keeps the running maximum;
//   - at the end of a tile, cnt = popcount(OR of all masks) gives inhibit's
//     step ts = T - max(cnt, 1); one MFMA pass over B fragments gathered
//     from the stage at each pixel's own ts yields the kept channel c*, and
//     a second at c*'s first firing step yields convwta's t0 value.  The
//     gathered passes issue the same MFMA sequence on the same operands as
//     the step loop, so they reproduce its potentials bit for bit.
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
    int32_t stage_bytes;  // bytes of the staged sample ([t][h][w][cin_pad] U8, 16-aligned)
    int32_t rows;         // window rows of a column tile (at most)
    int32_t pitch;        // 16-byte units per window row: cin_pad/8 planes of pw, padded to = wo (mod 16)
    int32_t win_units;    // 16-byte units of one step's FP16 window (rows * pitch)
    int32_t tc;           // steps whose windows are widened together
    int32_t stage_fast;   // staging by quad transposes (hw % 4 == 0, aligned input, cin_pad/4 | block quads)
    float th;
};

// Where K-group grp (8 input channels at one tap) of a pixel lives: byte
// offset in the stage relative to the pixel's own position, the tap's
// row/column offset (for the bounds test) and its unit offset in a window.
// Groups past the real K (zero weights) reuse group 0's operands.
struct TapEntry
{
    int32_t soff, dy, dx, woff;
};

__device__ __forceinline__ TapEntry tap_entry(const FusedGeom &g, int32_t grp)
{
    const int32_t cg = g.cin_pad / 8;
    const int32_t gg = grp < g.groups ? grp : 0, tap = gg / cg, c8 = gg % cg, ky = tap / g.kw, kx = tap % g.kw;
    const int32_t dy = ky - g.pt, dx = kx - g.pl;
    return TapEntry{(dy * g.w + dx) * g.cin_pad + c8 * 8, dy, dx, ky * g.pitch + c8 * g.pw + kx};
}


// Writes input byte i of the sample's [t][cin][h][w] tensor to its
// [t][y][x][c] slot of the (unpadded, channel-innermost) U8 stage.
__device__ __forceinline__ void stage_bytes(uint8_t *stage, const FusedGeom &g, int32_t i, uint8_t v, int32_t chw,
                                            int32_t hwi)
{
    const int32_t tt = i / chw, r = i - tt * chw, c = r / hwi, pix = r - c * hwi;
    stage[(tt * hwi + pix) * g.cin_pad + c] = v;
}

template <int S, bool DIAG>
__global__ void __launch_bounds__(512) fused_layer(const uint8_t *__restrict__ in, const half8 *__restrict__ frag,
                                                   Summary *__restrict__ summary, float *__restrict__ pot_out,
                                                   uint8_t *__restrict__ fs_out, FusedGeom g)
{
    extern __shared__ __align__(16) uint8_t lds[];
    uint8_t *stage = lds;                                                    // [t][h][w][cin_pad] U8
    uint4 *win16 = reinterpret_cast<uint4 *>(lds + g.stage_bytes);           // [tc][rows][pitch] FP16 x 8
    uint4 *patch = win16 + g.tc * g.win_units;                               // [4S][COLS] gathered B fragments
    uint32_t *any_s = reinterpret_cast<uint32_t *>(patch + 4 * S * COLS);    // [MAX_WAVES][COLS]
    float *red_v = reinterpret_cast<float *>(any_s + MAX_WAVES * COLS);      // [MAX_WAVES][COLS]
    int32_t *red_i = reinterpret_cast<int32_t *>(red_v + MAX_WAVES * COLS);  // [MAX_WAVES][COLS]
    int32_t *tsel = red_i + MAX_WAVES * COLS;                                // [COLS]
    int4 *gtab = reinterpret_cast<int4 *>(tsel + COLS);                      // [4S] TapEntry

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
    const int32_t hw = g.ho * g.wo, hwi = g.h * g.w, cg = g.cin_pad / 8;
    const int32_t o0 = w * 32 + 4 * q;  // the lane's first channel; rows r*16 + i follow

    // weight fragments and window offsets, resident for the whole kernel
    half8 a[2][S];
    int32_t offs[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        a[0][s] = frag[((w * 2 + 0) * S + s) * 64 + lane];
        a[1][s] = frag[((w * 2 + 1) * S + s) * 64 + lane];
        offs[s] = tap_entry(g, 4 * s + q).woff * 16;
    }
    for (int32_t e = threadIdx.x; e < 4 * S; e += blockDim.x)
    {
        const TapEntry te = tap_entry(g, e);
        gtab[e] = make_int4(te.soff, te.dy, te.dx, te.woff);
    }
    const int32_t win_step = g.win_units * 16;  // bytes of one step's window

    // window build: a thread owns one unit of a window row (channel plane
    // c8, padded column col) and walks the (step, row) pairs
    const int32_t row_units = cg * g.pw, lanes_per = int32_t(blockDim.x) / row_units;
    const int32_t cu = int32_t(threadIdx.x) % row_units, slot = int32_t(threadIdx.x) / row_units;
    const int32_t bc8 = cu / g.pw, bix = cu % g.pw - g.pl;
    const bool bcol = bix >= 0 && bix < g.w;

    for (int32_t n = blockIdx.x; n < g.n; n += gridDim.x)
    {
        __syncthreads();
        // stage the sample as [t][y][x][cin_pad] bytes
        if (g.stage_fast)
        {
            // a lane loads one dword (4 pixels of channel c) of the
            // [t][cin][h][w] input; the 4 lanes of a quad hold channels
            // 4*c4 .. 4*c4+3 of the same 4 pixels and transpose the 4x4 bytes
            // (two DPP exchanges + v_perm), after which lane l4 holds the 4
            // channels of pixel l4: one dword store per lane.  With
            // cin_pad / 4 dividing 16, the 64 stores of a wave hit 64 banks.
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
            // offset of the lane's dword, channel c of step tt
            int32_t lin = Q / D, pq = lin, off = (c * npq + pq) * 4;
            const int32_t total = g.t * npq, wrap_add = (g.cin - 1) * npq * 4;
            constexpr int U = 8;
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
                    const int32_t li = lin + u * step;
                    if (li < total) dst[(4 * li + l4) * D + c4] = __builtin_amdgcn_perm(y2, r1, sel2);
                }
            }
        }
        else
        {
            // any other shape: zero (for the padded channels), then scatter
            // the [t][cin][h][w] input byte by byte
            if (g.cin != g.cin_pad)
            {
                uint4 *z = reinterpret_cast<uint4 *>(stage);
                for (int32_t i = threadIdx.x; i < g.stage_bytes / 16; i += blockDim.x) z[i] = make_uint4(0, 0, 0, 0);
                __syncthreads();
            }
            const uint8_t *src = in + int64_t(n) * g.t * g.cin * hwi;
            const int32_t chw = g.cin * hwi, total = g.t * chw;
            for (int32_t i = int32_t(threadIdx.x); i < total; i += int32_t(blockDim.x))
                stage_bytes(stage, g, i, src[i], chw, hwi);
        }
        __syncthreads();
        if (fs_out)
        {
            // first spiking step of every staged input byte (t when it never
            // spikes), 8 channels of one pixel per thread: [n][h*w][cin_pad]
            const uint32_t tb = uint32_t(g.t) * 0x01010101u;
            for (int32_t u = threadIdx.x; u < hwi * cg; u += blockDim.x)
            {
                const uint8_t *sp = stage + u * 8;
                uint32_t lo = tb, hi = tb, seen_lo = 0, seen_hi = 0;
                for (int32_t tt = 0; tt < g.t; ++tt)
                {
                    const uint2 v = *reinterpret_cast<const uint2 *>(sp + tt * hwi * g.cin_pad);
                    first_nonzero(lo, seen_lo, v.x, tt);
                    first_nonzero(hi, seen_hi, v.y, tt);
                }
                *reinterpret_cast<uint2 *>(fs_out + (int64_t(n) * hwi * cg + u) * 8) = make_uint2(lo, hi);
            }
        }

        for (int32_t tile = 0; tile < g.tiles; ++tile)
        {
            const int32_t p = tile * COLS + j;
            const bool pvalid = p < hw;
            const int32_t pc = pvalid ? p : hw - 1, y = pc / g.wo, x = pc % g.wo;
            // the tile's window: padded rows [y_first, y_first + rows), all padded columns
            const int32_t y_first = (tile * COLS) / g.wo;
            const int32_t rows = min(tile * COLS + COLS - 1, hw - 1) / g.wo - y_first + g.kh;
            const uint8_t *wb = reinterpret_cast<const uint8_t *>(win16) + ((y - y_first) * g.pitch + x) * 16;
            uint32_t mask[2][4];
            int32_t mxb[2][4];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) mask[r][i] = 0, mxb[r][i] = 0;

            for (int32_t t0 = 0; t0 < g.t; t0 += g.tc)
            {
                const int32_t tc = min(g.tc, g.t - t0);
                // widen the windows of steps [t0, t0 + tc) once for the block:
                // a thread slot takes steps slot, slot + lanes_per, ... and walks
                // their rows; rows in the zero pad ([0, r_lo) and [r_hi, rows))
                // and columns outside the input are written 0
                if (slot < lanes_per)
                {
                    const int32_t r_lo = max(0, g.pt - y_first), r_hi = min(rows, g.h + g.pt - y_first);
                    const int32_t row_bytes = g.w * g.cin_pad;
                    for (int32_t ts = slot; ts < tc; ts += lanes_per)
                    {
                        const uint8_t *sp =
                            stage + (((t0 + ts) * g.h + y_first - g.pt) * g.w + bix) * g.cin_pad + bc8 * 8;
                        uint4 *dp = win16 + ts * g.win_units + cu;
                        for (int32_t r = 0; r < rows; ++r)
                        {
                            uint2 raw = make_uint2(0, 0);
                            if (bcol && r >= r_lo && r < r_hi)
                                raw = *reinterpret_cast<const uint2 *>(sp + r * row_bytes);
                            dp[r * g.pitch] = widen(raw);
                        }
                    }
                }
                __syncthreads();
                // the chunk's steps without barriers: the loads of step k+2 and
                // the MFMAs of step k+1 are issued before the update of step k
                // (loads past the chunk re-read its last step, unused)
                half8 b0[S], b1[S];
                float4v c0, c1, n0, n1;
                load_b(b0, wb, offs);
                mfma_chain(a, b0, c0, c1);
                load_b(b0, wb + min(1, tc - 1) * win_step, offs);
                auto update = [&](int32_t k, const float4v &u0, const float4v &u1) {
                    step_update(u0, u1, g.th, mask, mxb);
                    if (DIAG && pvalid)
#pragma unroll
                        for (int r = 0; r < 2; ++r)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                            {
                                const int32_t o = o0 + r * 16 + i;
                                if (o < g.cout)
                                    pot_out[((int64_t(n) * g.t + t0 + k) * g.cout + o) * hw + p] =
                                        r == 0 ? u0[i] : u1[i];
                            }
                };
                for (int32_t k = 0; k < tc; k += 2)
                {
                    load_b(b1, wb + min(k + 2, tc - 1) * win_step, offs);
                    if (k + 1 < tc) mfma_chain(a, b0, n0, n1);
                    update(k, c0, c1);
                    load_b(b0, wb + min(k + 3, tc - 1) * win_step, offs);
                    if (k + 2 < tc) mfma_chain(a, b1, c0, c1);
                    if (k + 1 < tc) update(k + 1, n0, n1);
                }
                if (t0 + tc < g.t) __syncthreads();  // windows free for the next chunk
            }

            // steps at which some channel fired, per pixel (bit T-1-t for step t)
            {
                uint32_t any = 0;
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) any |= mask[r][i];
                any |= __shfl_xor(any, 16, 64);
                any |= __shfl_xor(any, 32, 64);
                if (lane < COLS) any_s[w * COLS + lane] = any;
            }
            __syncthreads();
            // the B fragments of the tile's pixels, each at its own step st
            // (the same bytes the step loop's window held for that step)
            auto gather = [&](int32_t st) {
                for (int32_t grp = int32_t(threadIdx.x) >> 4; grp < 4 * S; grp += int32_t(blockDim.x) >> 4)
                {
                    const int4 te = gtab[grp];
                    const int32_t iy = y + te.y, ix = x + te.z;
                    uint2 raw = make_uint2(0, 0);
                    if (st >= 0 && iy >= 0 && iy < g.h && ix >= 0 && ix < g.w)
                        raw = *reinterpret_cast<const uint2 *>(stage + (st * hwi + y * g.w + x) * g.cin_pad + te.x);
                    patch[grp * COLS + j] = widen(raw);
                }
            };
            {
                uint32_t any = 0;
                for (int32_t ow = 0; ow < g.waves; ++ow) any |= any_s[ow * COLS + j];
                const int32_t cnt = __builtin_popcount(any);
                gather(g.t - (cnt > 1 ? cnt : 1));
            }
            __syncthreads();
            int32_t poffs[S];
#pragma unroll
            for (int s = 0; s < S; ++s) poffs[s] = (s * 64 + lane) * 16;
            // pass 1: the potentials at inhibit's step; the kept channel is the
            // first maximum of the thresholded values (inhibit.cpp:64-85).  The
            // lane's channels ascend with (r, i), so strict '>' keeps the first
            // maximum; padded channels have zero weights and value 0, which never
            // beats a lower channel.
            float4v g0, g1;
            {
                half8 b[S];
                load_b(b, reinterpret_cast<const uint8_t *>(patch), poffs);
                mfma_chain(a, b, g0, g1);
                float bv = -1.0f;
                int32_t bi = INT32_MAX;
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                    {
                        const float v = r == 0 ? g0[i] : g1[i];
                        const float xp = v > g.th ? v : 0.0f;
                        const bool take = xp > bv;
                        bv = take ? xp : bv;
                        bi = take ? o0 + r * 16 + i : bi;
                    }
                combine(bv, bi, __shfl_xor(bv, 16, 64), __shfl_xor(bi, 16, 64));
                combine(bv, bi, __shfl_xor(bv, 32, 64), __shfl_xor(bi, 32, 64));
                if (lane < COLS) red_v[w * COLS + lane] = bv, red_i[w * COLS + lane] = bi;
            }
            __syncthreads();
            int32_t cs;
            {
                float m = red_v[j];
                cs = red_i[j];
                for (int32_t ow = 1; ow < g.waves; ++ow) combine(m, cs, red_v[ow * COLS + j], red_i[ow * COLS + j]);
            }
            // the lane holding (c*, pixel) publishes c*'s first firing step
            const bool own = (cs >> 5) == w && ((cs >> 2) & 3) == q;
            uint32_t om = 0;
            int32_t ox = 0;
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (((cs >> 4) & 1) == r && (cs & 3) == i) om = mask[r][i], ox = mxb[r][i];
            if (own) tsel[j] = om ? g.t - 32 + int32_t(__builtin_clz(om)) : -1;
            __syncthreads();
            gather(tsel[j]);
            __syncthreads();
            // pass 2: the value of c* at its first firing step (convwta's t0 value)
            float first = 0.0f;
            if (__builtin_amdgcn_ballot_w64(own && om != 0))
            {
                half8 b[S];
                load_b(b, reinterpret_cast<const uint8_t *>(patch), poffs);
                mfma_chain(a, b, g0, g1);
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (((cs >> 4) & 1) == r && (cs & 3) == i) first = r == 0 ? g0[i] : g1[i];
            }
            if (own && pvalid)
            {
                const uint32_t fwd = om ? __builtin_bitreverse32(om) >> (32 - g.t) : 0u;
                summary[int64_t(n) * hw + p] =
                    Summary{cs, fwd, om ? first : 0.0f, om ? __builtin_bit_cast(float, ox) : 0.0f};
            }
        }
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
    for (int32_t r = k + int32_t(threadIdx.x); r < count; r += blockDim.x) wkey[s * count + r] = -1;
}

// Spikes [n][t][cout][ho][wo] and (optionally) their max-pool, from the
// summaries.  One block per (sample, step); each thread writes 16 bytes.
__global__ void __launch_bounds__(256) fused_write(const Summary *__restrict__ summary, uint8_t *__restrict__ out,
                                                   uint8_t *__restrict__ pooled, int32_t t, int32_t cout, int32_t ho,
                                                   int32_t wo, int32_t pkh, int32_t pkw, int32_t psh, int32_t psw,
                                                   int32_t pho, int32_t pwo)
{
    extern __shared__ __align__(16) uint8_t lds[];
    int16_t *act = reinterpret_cast<int16_t *>(lds);  // active channel per pixel at this step, or -1
    const int32_t hw = ho * wo;
    const int64_t ns = blockIdx.x / t;
    const int32_t step = blockIdx.x % t;
    for (int32_t p = threadIdx.x; p < hw; p += blockDim.x)
    {
        const Summary e = summary[ns * hw + p];
        act[p] = (e.mask >> step) & 1u ? int16_t(e.c) : int16_t(-1);
    }
    __syncthreads();
    const int64_t total_out = int64_t(cout) * hw;
    if (out && hw % 4 == 0 && (reinterpret_cast<uintptr_t>(out + int64_t(blockIdx.x) * total_out) & 3) == 0)
    {
        // four pixels of one channel plane per 32-bit store, consecutive
        // threads on consecutive pixels of the same plane
        uint8_t *dst = out + int64_t(blockIdx.x) * total_out;
        const int32_t quads = hw / 4, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
        for (int32_t qd = lane; qd < quads; qd += 64)
        {
            // the lane's four pixels keep their active channels in registers
            // while the wave walks the channel planes
            const uint2 a2 = *reinterpret_cast<const uint2 *>(act + 4 * qd);
            const int32_t a0 = int16_t(a2.x & 0xFFFF), a1 = int16_t(a2.x >> 16);
            const int32_t a2v = int16_t(a2.y & 0xFFFF), a3v = int16_t(a2.y >> 16);
            uint8_t *d = dst + 4 * qd;
            for (int32_t c = wave; c < cout; c += waves)
            {
                const uint32_t word = uint32_t(a0 == c) | uint32_t(a1 == c) << 8 | uint32_t(a2v == c) << 16 |
                                      uint32_t(a3v == c) << 24;
                *reinterpret_cast<uint32_t *>(d + int64_t(c) * hw) = word;
            }
        }
    }
    else if (out)
    {
        // bytes [g0, g0 + total) of out, walked in 16-byte aligned chunks
        const int64_t total = total_out;
        const int64_t g0 = int64_t(blockIdx.x) * total;
        const int64_t head = int64_t(reinterpret_cast<uintptr_t>(out + g0) & 15);
        for (int64_t c0 = int64_t(threadIdx.x) * 16 - head; c0 < total; c0 += int64_t(blockDim.x) * 16)
        {
            uint8_t v[16];
            const int64_t start = c0 < 0 ? 0 : c0;
            int32_t c = int32_t(start / hw), p = int32_t(start % hw);
#pragma unroll
            for (int b = 0; b < 16; ++b)
            {
                const int64_t i = c0 + b;
                v[b] = (i >= 0 && i < total && act[p] == c) ? 1 : 0;
                if (i >= 0 && ++p == hw) p = 0, ++c;
            }
            uint8_t *d = out + g0 + c0;
            if (c0 >= 0 && c0 + 16 <= total)
                *reinterpret_cast<uint4 *>(d) = *reinterpret_cast<uint4 *>(v);
            else
                for (int b = 0; b < 16; ++b)
                    if (c0 + b >= 0 && c0 + b < total) d[b] = v[b];
        }
    }
    if (pooled)
    {
        const int32_t phw = pho * pwo, total = cout * phw, win = pkh * pkw;
        uint8_t *dst = pooled + int64_t(blockIdx.x) * total;
        if (win <= 16)
        {
            // each thread owns one pooled position (its window's active
            // channels in registers) and a strided set of channels
            const int32_t P = min(phw, int32_t(blockDim.x)), C = int32_t(blockDim.x) / P;
            const int32_t clane = int32_t(threadIdx.x) / P;
            for (int32_t pos = int32_t(threadIdx.x) % P; clane < C && pos < phw; pos += P)
            {
                const int32_t py = pos / pwo, px = pos % pwo;
                int32_t wa[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                {
                    const int32_t y = py * psh + k / pkw, x = px * psw + k % pkw;
                    wa[k] = (k < win && y < ho && x < wo) ? act[y * wo + x] : -1;
                }
                for (int32_t c = clane; c < cout; c += C)
                {
                    bool m = false;
#pragma unroll
                    for (int k = 0; k < 16; ++k) m |= wa[k] == c;
                    dst[int64_t(c) * phw + pos] = m;
                }
            }
        }
        else
            for (int32_t i = threadIdx.x; i < total; i += blockDim.x)
            {
                const int32_t c = i / phw, r = i % phw, py = r / pwo, px = r % pwo;
                uint8_t m = 0;
                for (int32_t a = 0; a < pkh; ++a)
                    for (int32_t b = 0; b < pkw; ++b)
                    {
                        const int32_t y = py * psh + a, x = px * psw + b;
                        if (y < ho && x < wo && act[y * wo + x] == c) m = 1;
                    }
                dst[i] = m;
            }
    }
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
    // one non-blocking side stream per device, at the default priority
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lock(mu);
        if (int(streams.size()) <= dev) streams.resize(size_t(dev) + 1, nullptr);
        if (!streams[size_t(dev)]) SPK_HIP(hipStreamCreateWithFlags(&streams[size_t(dev)], hipStreamNonBlocking));
        stream = streams[size_t(dev)];
    }
    hipEvent_t fork;
    SPK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    SPK_HIP(hipEventRecord(fork, main));
    SPK_HIP(hipStreamWaitEvent(stream, fork, 0));
    SPK_HIP(hipEventDestroy(fork));
}

void SideStream::join()
{
    if (stream == main || joined) return;
    hipEvent_t done;
    SPK_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    SPK_HIP(hipEventRecord(done, stream));
    SPK_HIP(hipStreamWaitEvent(main, done, 0));
    SPK_HIP(hipEventDestroy(done));
    joined = true;
}

SideStream::~SideStream()
{
    if (stream != main && !joined)
    {
        // an error path: still order the call's stream after the side work
        hipEvent_t done;
        if (hipEventCreateWithFlags(&done, hipEventDisableTiming) == hipSuccess)
        {
            (void)hipEventRecord(done, stream);
            (void)hipStreamWaitEvent(main, done, 0);
            (void)hipEventDestroy(done);
        }
    }
}

namespace
{
template <int S, bool DIAG>
void launch_layer(const uint8_t *in, const half8 *frag, Summary *summary, float *pot, uint8_t *fs, const FusedGeom &g,
                  size_t lds,
                  hipStream_t s)
{
    static bool configured = false;  // per instantiation; the attribute is per function
    if (!configured)
    {
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_layer<S, DIAG>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds > 65536 ? 163840 : 65536)));
        configured = true;
    }
    int dev = 0, cus = 0;
    SPK_HIP(hipGetDevice(&dev));
    SPK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = g.n < cus ? g.n : cus;
    KernelTimer &tm = timer();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tm.on)
    {
        SPK_HIP(hipEventCreate(&e0));
        SPK_HIP(hipEventCreate(&e1));
        SPK_HIP(hipEventRecord(e0, s));
    }
    fused_layer<S, DIAG><<<dim3(grid), dim3(64 * g.waves), lds, s>>>(in, frag, summary, pot, fs, g);
    SPK_LAUNCHED();
    if (tm.on)
    {
        SPK_HIP(hipEventRecord(e1, s));
        tm.events.emplace_back(e0, e1);
    }
}

void dispatch_layer(int steps, const uint8_t *in, const half8 *frag, Summary *summary, float *pot, uint8_t *fs,
                    const FusedGeom &g, size_t lds, hipStream_t s)
{
    switch (steps)
    {
#define CASE(k) \
    case k:                                                                                 \
        return pot ? launch_layer<k, true>(in, frag, summary, pot, fs, g, lds, s)             \
                   : launch_layer<k, false>(in, frag, summary, pot, fs, g, lds, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: SPK_ARG(false, "Fused layer: K-steps out of range.");
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

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
        SPK_ARG(n >= 1 && t >= 1 && cin >= 1 && h >= 1 && w >= 1 && cout >= 2 && kh >= 1 && kw >= 1,
                "Fused layer: invalid shapes.");
        SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
        SPK_ARG(t <= 32, "Fused layer: at most 32 time steps.");
        SPK_ARG(cout <= 32 * MAX_WAVES, "Fused layer: at most " << 32 * MAX_WAVES << " output channels.");
        SPK_ARG(threshold >= 0.0, "Fused layer: the conv threshold must be >= 0.");
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
        g.waves = int32_t(cdiv(cout, 32));
        g.tiles = int32_t(cdiv(ho * wo, COLS));
        g.stage_bytes = int32_t(cdiv(int64_t(t) * h * w * g.cin_pad, 16) * 16);
        // a tile of COLS row-major pixels spans at most ceil((COLS-1)/wo)+1 rows
        g.rows = int32_t(cdiv(COLS - 1, wo) + kh);
        // row pitch = wo (mod 16) puts 16 consecutive output pixels in 16
        // distinct 16-byte bank groups
        const int64_t row_units = int64_t(g.pw) * (g.cin_pad / 8);
        g.pitch = int32_t(row_units + ((wo - row_units) % 16 + 16) % 16);
        g.win_units = g.rows * g.pitch;
        g.th = float(threshold);
        SPK_ARG(g.steps <= MAX_STEPS, "Fused layer: Cin*Kh*Kw too large (K <= 512 after padding).");
        // stage, gathered fragments, any/argmax partials, step table, tap table
        const size_t fixed = size_t(g.stage_bytes) + size_t(g.steps) * 4 * COLS * 16 + MAX_WAVES * COLS * 12 +
                             COLS * 4 + size_t(g.steps) * 4 * 16;
        const int64_t room = (163840 - int64_t(fixed)) / (int64_t(g.win_units) * 16);
        if (room < 1 || row_units > 64 * g.waves || ho * wo > 8192 || fixed > 163840)
        {
            // the sample does not fit the LDS whole: the region path
            const StdpConfigs rcfg = learn ? pack_configs(configs, nconfigs) : StdpConfigs{};
            region_layer_step((const uint8_t *)in, n, t, cin, h, w, (float *)kernel, cout, kh, kw, pt, pl, pb, pr,
                              threshold, ry, rx, count, learn ? &rcfg.c[0] : nullptr, (uint8_t *)out,
                              (uint8_t *)pooled, pkh, pkw, psh, psw, winners, counts, (float *)potentials, S(stream));
            return;
        }
        const int64_t chunks = cdiv(t, room < t ? room : t);
        g.tc = int32_t(cdiv(t, chunks));
        g.stage_fast = (h * w) % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 3) == 0 &&
                       (16 * g.waves) % (g.cin_pad / 4) == 0 && 16 * g.waves / (g.cin_pad / 4) <= h * w / 4;
        const size_t lds = fixed + size_t(g.tc) * g.win_units * 16;
        SPK_ARG(lds <= 163840, "Fused layer: the staged sample (" << lds << " bytes) exceeds the 160 KiB LDS.");
        SPK_ARG(int64_t(cout) * ho * wo < (1LL << 31), "Fused layer: output plane too large.");

        const StdpConfigs cfg = learn ? pack_configs(configs, nconfigs) : StdpConfigs{};
        SPK_ARG(!learn || n * count < (1LL << 31), "Fused layer: too many winner slots.");

        hipStream_t s = S(stream);
        const int64_t hw = ho * wo, hwi = h * w;
        Scratch frag(int64_t(g.waves) * 2 * g.steps * 64 * 16, s), summary(n * hw * int64_t(sizeof(Summary)), s);
        Scratch wkey(n * count * int64_t(sizeof(int32_t)), s), fsm(learn ? n * hwi * g.cin_pad : 0, s);
        prep_fragments<<<dim3(unsigned(cdiv(int64_t(g.waves) * 2 * g.steps * 64, 256))), 256, 0, s>>>(
            (const float *)kernel, frag.as<half8>(), g.cout, g.cin, g.kh, g.kw, g.cin_pad, g.groups, g.steps,
            g.waves * 2);
        SPK_LAUNCHED();
        dispatch_layer(g.steps, (const uint8_t *)in, frag.as<half8>(), summary.as<Summary>(), (float *)potentials,
                       learn ? fsm.as<uint8_t>() : nullptr, g, lds, s);
        // the writer needs only the summaries: it runs on a side stream,
        // overlapping convwta, the records and the STDP chains (HBM-bound
        // beside latency-bound), and the call's stream waits for it at the end
        SideStream side(s, out || pooled);
        if (out || pooled)
        {
            fused_write<<<dim3(unsigned(n * t)), 256, size_t(hw) * 2, side.stream>>>(
                summary.as<Summary>(), (uint8_t *)out, (uint8_t *)pooled, g.t, g.cout, g.ho, g.wo, int32_t(pkh),
                int32_t(pkw), int32_t(psh), int32_t(psw), int32_t(pho), int32_t(pwo));
            SPK_LAUNCHED();
        }
        fused_select<<<dim3(unsigned(n)), 256, size_t(hw) * 8, s>>>(summary.as<Summary>(), g.t, g.ho, g.wo,
                                                                    int32_t(ry), int32_t(rx), int32_t(count), winners,
                                                                    counts, wkey.as<int32_t>());
        SPK_LAUNCHED();
        if (learn)
        {
            const int64_t ksz = cin * kh * kw;
            const FusedStdp fp{n * count, int32_t(count), g.cin, g.cin_pad, g.h, g.w, g.kh, g.kw, g.pt, g.pl,
                               g.wo, int32_t(hw), g.t, int32_t(cdiv(ksz, 32))};
            Scratch bits(n * count * fp.words * int64_t(sizeof(uint32_t)), s);
            fused_records<<<dim3(unsigned(cdiv(n * count, 4))), 256, 0, s>>>(
                fsm.as<uint8_t>(), summary.as<Summary>(), wkey.as<int32_t>(), fp, bits.as<uint32_t>());
            SPK_LAUNCHED();
            fused_stdp<<<dim3(unsigned(cout), unsigned(cdiv(ksz, STDP_THREADS))), STDP_THREADS, 0, s>>>(
                (float *)kernel, bits.as<uint32_t>(), wkey.as<int32_t>(), fp, cfg.c[0]);
            SPK_LAUNCHED();
        }
        side.join();
    });
}
}
