// Fused FP16-MFMA layer step over output regions: for samples whose input
// does not fit the LDS whole, and for batches of instances of uneven size and
// step count (one persistent launch per layer).
//
// The semantics are those of the whole-sample kernel (fused.hip): conv ->
// threshold -> inhibit -> fire summary per output pixel, the spikes and pooled
// spikes, then convwta and STDP on the FP32 masters, in the reference's batch
// order (src/spyker/opers.cpp:408-693).  What changes is the unit of work:
//
//   - a work item is (instance, region), a region being RH output rows x 16
//     output columns; the LDS holds the region's input rectangle with its halo
//     for all of the instance's steps, [t][SH][SW][cin_pad] U8, read from the
//     instance's [t][cin][h][w] input (dword loads and quad transposes when
//     w % 4 == 0);
//   - a block of CG x NCT waves works on NCT region rows at once: wave
//     (channel group, column tile) owns 32 output channels of one row of 16
//     pixels, its FP16 weight fragments in VGPRs for the whole kernel;
//   - the FP16 windows of a chunk of steps cover the NCT rows and their halo;
//     every wave runs the chunk's steps without barriers, as in fused.hip;
//   - at a row's end the gathered MFMA passes (inhibit's kept channel, then
//     convwta's t0 value) load each lane's B operands straight from the
//     stage at the pixel's own step, the same bytes and the same MFMA
//     sequence as the step loop, so they reproduce its potentials exactly;
//   - the region also writes the first-spike map of the input pixels it owns
//     (a partition of the input), for the STDP records;
//   - the spike block of a group of rows is zeroed by 16-byte non-temporal
//     stores spread over its step loop (and the pooled block when the pool
//     windows lie inside the group), and the group's end stores the kept
//     channels' 1s over them: at most one 1 per pixel and step.
//
// A plan (spk_layer_plan_*) fixes the shapes once: instance and item tables
// and every work buffer are allocated when it is made, so running it
// allocates nothing and never waits on the host, and a stream capture
// (hipGraph) of a run replays it.  Region height and chunking are chosen by
// timing the candidates on the first run when the plan was made with tuning.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "fused.hpp"

namespace spk
{
namespace
{
using u32x4 = __attribute__((ext_vector_type(4))) uint32_t;
constexpr int RW = 16;            // region width: one column tile per region row
constexpr int REGION_WAVES = 8;   // waves per block
constexpr int LDS_BYTES = 163840;
constexpr int SU = 16;  // stage rows a thread loads together
constexpr int WU = 4;  // window (step, row) pairs a thread converts together

struct RegionInst
{
    int64_t in_off;    // bytes: the instance's input [t][cin][h][w]
    int64_t out_off;   // bytes: its spikes [t][cout][ho][wo]
    int64_t pool_off;  // bytes: its pooled spikes [t][cout][pho][pwo]
    int64_t sum_off;   // its first Summary (and select slot)
    int64_t fs_off;    // bytes: its first-spike map [h*w][cin_pad]
    int32_t t, h, w, ho, wo, pho, pwo, cols;  // cols = regions per region row
};

struct RegionGeom
{
    int32_t cin, cout, kh, kw, pt, pl;
    int32_t cin_pad, groups, steps, cg, nct, waves;
    int32_t rh;                  // region rows, a multiple of nct
    int32_t sh, sw;              // stage rows and columns (sw a multiple of 4)
    int32_t wrows, wcols, pitch, win_units, tc;
    int32_t stage_bytes;         // tmax * sh * sw * cin_pad, 16-aligned
    int32_t fast;                // dword staging (every w % 4 == 0, 4-aligned inputs)
    int32_t pair;                // cin <= 4: K groups of two taps x 4 channels, 8-byte window units
    int64_t items;
    float th;
};

// The step's outputs the region kernel writes itself: spikes [t][cout][ho][wo]
// and, when every pool window lies inside one group of region rows
// (non-overlapping windows, nct a multiple of pkh, 16 a multiple of pkw),
// pooled spikes [t][cout][pho][pwo].  Either pointer may be null.
struct RegionOut
{
    uint8_t *out;
    uint8_t *pooled;
    int32_t pkh, pkw;
    // [instance] max(0, every pixel's maximum), as FP32 bits (their unsigned
    // order is the order of non-negative floats); the caller zeroes it
    uint32_t *imax;
};

// cin <= 4: K groups of two taps x 4 channels when that needs at most 8 K-steps.
inline bool use_pair(int64_t cin, int64_t kh, int64_t kw) { return cin <= 4 && cdiv(kh * ((kw + 1) / 2), 4) <= 8; }

// Weight fragments of the paired K groups: lane l of K step s holds output
// channel rt*16 + l%16 and group 4s + l/16 = (ky, p); half j is channel j%4
// of tap (ky, 2p + j/4), 0 past the kernel.
__global__ void prep_fragments_pair(const float *__restrict__ k, half8 *__restrict__ frag, int32_t cout, int32_t cin,
                                    int32_t kh, int32_t kw, int32_t groups, int32_t steps, int32_t row_tiles)
{
    const int32_t total = row_tiles * steps * 64, np = (kw + 1) / 2;
    for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x)
    {
        const int32_t lane = e % 64, s = (e / 64) % steps, rt = e / (64 * steps);
        const int32_t o = rt * 16 + lane % 16, grp = 4 * s + lane / 16, ky = grp / np, p = grp % np;
        half8 v;
        for (int j = 0; j < 8; ++j)
        {
            const int32_t kx = 2 * p + j / 4, c = j % 4;
            float x = 0.0f;
            if (o < cout && grp < groups && c < cin && kx < kw) x = k[((int64_t(o) * cin + c) * kh + ky) * kw + kx];
            v[j] = _Float16(x);
        }
        frag[e] = v;
    }
}

// LDS bytes past the stage and windows: any/argmax partials, step table, tap table.
inline int64_t region_fixed(int32_t waves, int32_t nct, int32_t steps)
{
    return int64_t(waves) * COLS * 12 + int64_t(nct) * COLS * 12 + int64_t(steps) * 4 * 16;
}

template <int S, bool DIAG, bool PAIR, bool BSTAGE>
__global__ void __launch_bounds__(64 * REGION_WAVES)
    fused_region(const uint8_t *__restrict__ in, const half8 *__restrict__ frag, const RegionInst *__restrict__ inst,
                 const int2 *__restrict__ items, Summary *__restrict__ summary, float *__restrict__ pot_out,
                 uint8_t *__restrict__ fs_out, RegionOut ro, RegionGeom g)
{
    extern __shared__ __align__(16) uint8_t lds[];
    uint8_t *stage = lds;                                                          // [t][sh][sw][cin_pad]
    uint4 *win16 = reinterpret_cast<uint4 *>(lds + g.stage_bytes);                 // [tc][wrows][pitch]
    uint32_t *any_s = reinterpret_cast<uint32_t *>(win16 + g.tc * g.win_units);    // [waves][COLS]
    float *red_v = reinterpret_cast<float *>(any_s + g.waves * COLS);             // [waves][COLS]
    int32_t *red_i = reinterpret_cast<int32_t *>(red_v + g.waves * COLS);         // [waves][COLS]
    int32_t *tsel = red_i + g.waves * COLS;                                       // [nct][COLS]
    int32_t *sel_c = tsel + g.nct * COLS;                                         // [nct][COLS] kept channel
    uint32_t *sel_m = reinterpret_cast<uint32_t *>(sel_c + g.nct * COLS);         // [nct][COLS] its steps
    int4 *gtab = reinterpret_cast<int4 *>(sel_m + g.nct * COLS);                  // [4S] (dy, dx, byte, 0)

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
    const int32_t cgw = w % g.cg, ct = w / g.cg;
    const int32_t o0 = cgw * 32 + 4 * q;  // the lane's first channel; rows r*16 + i follow
    // K groups: 8 input channels at one tap, or (PAIR, cin <= 4) the 4
    // channels of taps (ky, 2p) and (ky, 2p + 1), whose window units (8 bytes
    // a pixel) are adjacent
    const int32_t planes = PAIR ? 1 : g.cin_pad / 8, np = (g.kw + 1) / 2;
    constexpr int32_t UB = PAIR ? 8 : 16;  // bytes of one window unit

    // weight fragments and window offsets, resident for the whole kernel;
    // groups past the real K (zero weights) reuse group 0's operands
    half8 a[2][S];
    int32_t offs[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        a[0][s] = frag[((cgw * 2 + 0) * S + s) * 64 + lane];
        a[1][s] = frag[((cgw * 2 + 1) * S + s) * 64 + lane];
        const int32_t grp = 4 * s + q, gg = grp < g.groups ? grp : 0;
        if (PAIR)
            offs[s] = ((gg / np) * g.pitch + 2 * (gg % np)) * UB;
        else
        {
            const int32_t tap = gg / planes, c8 = gg % planes;
            offs[s] = ((tap / g.kw) * g.pitch + c8 * g.wcols + tap % g.kw) * UB;
        }
    }
    for (int32_t e = threadIdx.x; e < 4 * S; e += blockDim.x)
    {
        const int32_t gg = e < g.groups ? e : 0;
        if (PAIR)
            gtab[e] = make_int4(gg / np - g.pt, 2 * (gg % np) - g.pl, 0, 0);
        else
        {
            const int32_t tap = gg / planes, c8 = gg % planes;
            gtab[e] = make_int4(tap / g.kw - g.pt, tap % g.kw - g.pl, c8 * 8, 0);
        }
    }
    const int32_t win_step = g.win_units * 16;

    // window build roles: a thread owns one unit (plane, column) of a window
    // row and walks the (step, row) pairs of a chunk by `wslots`
    const int32_t row_units = planes * g.wcols, wslots = int32_t(blockDim.x) / row_units;
    const int32_t wcu = int32_t(threadIdx.x) % row_units, wslot = int32_t(threadIdx.x) / row_units;
    const int32_t wc8 = wcu / g.wcols, wcc = wcu % g.wcols;
    const int32_t wdq = wslots / g.wrows, wdr = wslots % g.wrows;

    // staging roles (dword path): quad Q of a slot holds channel quad c4 of
    // pixel quad lq; the slot walks the (step, stage row) pairs
    const int32_t D = g.cin_pad / 4, nq = g.sw / 4, sroles = nq * D * 4;
    const int32_t sslot = int32_t(threadIdx.x) / sroles, srole = int32_t(threadIdx.x) % sroles;
    const int32_t sslots = int32_t(blockDim.x) / sroles;
    const int32_t sQ = srole >> 2, l4 = srole & 3, c4 = sQ % D, lq = sQ / D, sc = 4 * c4 + l4;
    const uint32_t sel1 = (l4 & 1) ? 0x03070105u : 0x06020400u;
    const uint32_t sel2 = (l4 & 2) ? 0x03020706u : 0x05040100u;
    // staging roles (byte path): (channel, column) pairs, columns fastest so
    // a wave reads consecutive input bytes; a thread takes every
    // blockDim-th pair and walks the (step, stage row) pairs
    const int32_t broles = g.sw * g.cin;

    // output roles: each step of a group's step loop, thread z < cout * nct
    // zeroes the 16 spike bytes of (channel, row) z, and the next cout *
    // nct / pkh threads the pooled bytes of (channel, pooled row); the
    // group's end stores the 1s over them.  A spike block is all zero but for
    // at most one 1 per pixel and step (its kept channel), so the zero stores
    // stream out between the MFMAs and only the 1s remain for the end.
    const int32_t zs = ro.out ? g.cout * g.nct : 0;
    const int32_t prows = ro.pooled ? g.nct / ro.pkh : 0, zp = g.cout * prows;
    const int32_t zrole = int32_t(threadIdx.x);
    const bool zspk = zrole < zs, zpool = zrole >= zs && zrole < zs + zp;
    const int32_t zc = zspk ? zrole / g.nct : (zrole - zs) / max(prows, 1);
    const int32_t zr = zspk ? zrole % g.nct : (zrole - zs) % max(prows, 1);
    const int32_t pbytes = RW / max(ro.pkw, 1);  // pooled bytes of a region row

    // blocks of one XCD (dispatched round-robin over 8) take consecutive
    // items, so neighbouring regions share their halo in that XCD's L2
    const int64_t nb = gridDim.x;
    const int64_t first = nb % 8 == 0 ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;

    // the wave's maximum of its instance's pixels so far (the instance's
    // maximum for convwta, -0 and below counting as +0), added to the
    // instance's entry with one atomicMax when the block moves to another
    // instance and at the end, rather than one atomic a group
    float imx = 0.0f;
    int32_t imx_inst = -1;
    auto flush_imax = [&]() {
        if (lane == 0 && imx_inst >= 0 && imx > 0.0f)
            atomicMax(ro.imax + imx_inst, __builtin_bit_cast(uint32_t, imx));
    };

    for (int64_t it = first; it < g.items; it += nb)
    {
        const int2 itm = items[it];
        const RegionInst I = inst[itm.x];
        const int32_t oy0 = (itm.y / I.cols) * g.rh, ox0 = (itm.y % I.cols) * RW;
        const int32_t sy0 = oy0 - g.pt, sx0 = (ox0 - g.pl) & ~3;  // stage origin (input coordinates)
        const int32_t xoff = ox0 - g.pl - sx0;
        const uint8_t *src = in + I.in_off;
        uint8_t *const out_i = ro.out ? ro.out + I.out_off : nullptr;
        uint8_t *const pool_i = ro.pooled ? ro.pooled + I.pool_off : nullptr;
        const bool spk16 = out_i && I.wo % 16 == 0 && (reinterpret_cast<uintptr_t>(out_i) & 15) == 0;
        const bool pool8 = pool_i && pbytes == 8 && I.pwo % 8 == 0 && (reinterpret_cast<uintptr_t>(pool_i) & 7) == 0;
        __syncthreads();  // the previous item's stage and windows are free

        // stage the input rectangle [sy0, sy0 + sh) x [sx0, sx0 + sw), all steps
        const int32_t srows = I.t * g.sh;
        if (g.fast)
        {
            if (sslot < sslots)
            {
                uint32_t *dst = reinterpret_cast<uint32_t *>(stage);
                const int32_t ix0 = sx0 + 4 * lq;
                const bool colv = ix0 >= 0 && ix0 < I.w && sc < g.cin;
                if constexpr (BSTAGE)
                {
                    // raw buffer loads over the instance's input (fewer than
                    // 2^31 bytes, checked when the plan is made) at an offset
                    // kept by additions (step tt, channel sc, input row
                    // sy0 + lr, column ix0) as the slot walks the flattened
                    // stage rows r = tt * sh + lr by sslots = dq * sh + dr;
                    // the offset of its end reads 0, which stands for rows
                    // above or below the image, columns or channels outside it
                    // and the walk's tail
                    const int32_t plane = I.h * I.w, tstride = g.cin * plane, bytes = I.t * tstride;
                    const __amdgpu_buffer_rsrc_t rs =
                        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(src), 0, bytes, 0x00020000);
                    const int32_t dq = sslots / g.sh, dr = sslots % g.sh;
                    int32_t lr = sslot % g.sh;
                    int32_t off = (sslot / g.sh) * tstride + sc * plane + (sy0 + lr) * I.w + ix0;
                    const int32_t d_off = dq * tstride + dr * I.w, wrap_off = tstride - g.sh * I.w;
                    const int32_t lane_at = (4 * lq + l4) * D + c4, row_at = g.sw * D;
                    for (int32_t r0 = sslot; r0 < srows; r0 += SU * sslots)
                    {
                        uint32_t v[SU];
#pragma unroll
                        for (int u = 0; u < SU; ++u)
                        {
                            const int32_t iy = sy0 + lr;
                            const bool ok = r0 + u * sslots < srows && colv && iy >= 0 && iy < I.h;
                            v[u] = uint32_t(__builtin_amdgcn_raw_buffer_load_b32(rs, ok ? off : bytes, 0, 0));
                            lr += dr, off += d_off;
                            if (lr >= g.sh) lr -= g.sh, off += wrap_off;
                        }
#pragma unroll
                        for (int u = 0; u < SU; ++u)
                        {
                            // the 4x4 byte transpose of the flat walk below
                            const uint32_t x = v[u];
                            const uint32_t y = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(x), 0xB1, 0xF, 0xF, true));
                            const uint32_t r1 = __builtin_amdgcn_perm(y, x, sel1);
                            const uint32_t y2 = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(r1), 0x4E, 0xF, 0xF, true));
                            if (r0 + u * sslots < srows)
                                dst[(r0 + u * sslots) * row_at + lane_at] = __builtin_amdgcn_perm(y2, r1, sel2);
                        }
                    }
                }
                else
                {
                    int32_t tt = sslot / g.sh, lr = sslot % g.sh;
                    const int32_t dq = sslots / g.sh, dr = sslots % g.sh;
                    // batches of SU rows a slot, the batch's loads in flight together
                    for (int32_t r0 = sslot; r0 < srows; r0 += SU * sslots)
                    {
                        uint32_t v[SU];
                        int32_t at[SU];
#pragma unroll
                        for (int u = 0; u < SU; ++u)
                        {
                            const int32_t iy = sy0 + lr;
                            v[u] = 0;
                            if (r0 + u * sslots < srows && colv && iy >= 0 && iy < I.h)
                                v[u] = *reinterpret_cast<const uint32_t *>(
                                    src + ((int64_t(tt) * g.cin + sc) * I.h + iy) * I.w + ix0);
                            at[u] = ((tt * g.sh + lr) * g.sw + 4 * lq + l4) * D + c4;
                            tt += dq, lr += dr;
                            if (lr >= g.sh) lr -= g.sh, ++tt;
                        }
#pragma unroll
                        for (int u = 0; u < SU; ++u)
                        {
                            // 4x4 byte transpose within the quad: lane l4 then holds
                            // channels 4*c4 .. 4*c4+3 of pixel 4*lq + l4
                            const uint32_t x = v[u];
                            const uint32_t y = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(x), 0xB1, 0xF, 0xF, true));
                            const uint32_t r1 = __builtin_amdgcn_perm(y, x, sel1);
                            const uint32_t y2 = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(r1), 0x4E, 0xF, 0xF, true));
                            if (r0 + u * sslots < srows) dst[at[u]] = __builtin_amdgcn_perm(y2, r1, sel2);
                        }
                    }
                }
            }
        }
        else
        {
            // the padded channels cin .. cin_pad of every staged pixel are 0
            if (g.cin != g.cin_pad)
                for (int32_t u = threadIdx.x; u < srows * g.sw; u += blockDim.x)
                    for (int32_t c = g.cin; c < g.cin_pad; ++c) stage[u * g.cin_pad + c] = 0;
            for (int32_t role = threadIdx.x; role < broles; role += blockDim.x)
            {
                const int32_t blc = role % g.sw, bc = role / g.sw, ix = sx0 + blc;
                const bool colv = ix >= 0 && ix < I.w;
                int32_t tt = 0, lr = 0;
                for (int32_t r = 0; r < srows; ++r)
                {
                    const int32_t iy = sy0 + lr;
                    uint8_t v = 0;
                    if (colv && iy >= 0 && iy < I.h) v = src[((int64_t(tt) * g.cin + bc) * I.h + iy) * I.w + ix];
                    stage[((tt * g.sh + lr) * g.sw + blc) * g.cin_pad + bc] = v;
                    if (++lr == g.sh) lr = 0, ++tt;
                }
            }
        }
        __syncthreads();

        if (fs_out)
        {
            // first spiking step of the input pixels this region owns (t when
            // never): rows [oy0 - pt, oy0 + rh - pt), columns [ox0 - pl,
            // ox0 + 16 - pl), clipped, the last band of either extending to
            // the input's edge; 8 channels per thread, [h*w][cin_pad]
            const int32_t iy_lo = max(0, sy0), ix_lo = max(0, ox0 - g.pl);
            const int32_t iy_hi = oy0 + g.rh >= I.ho ? I.h : min(I.h, oy0 + g.rh - g.pt);
            const int32_t ix_hi = ox0 + RW >= I.wo ? I.w : min(I.w, ox0 + RW - g.pl);
            const int32_t ow = ix_hi - ix_lo, units = (iy_hi - iy_lo) * ow * planes;
            const uint32_t tb = uint32_t(I.t) * 0x01010101u;
            for (int32_t u = threadIdx.x; u < units; u += blockDim.x)
            {
                const int32_t c8 = u % planes, px = u / planes, iy = iy_lo + px / ow, ix = ix_lo + px % ow;
                const uint8_t *sp = stage + ((iy - sy0) * g.sw + ix - sx0) * g.cin_pad + c8 * 8;
                uint8_t *fp = fs_out + I.fs_off + (int64_t(iy) * I.w + ix) * g.cin_pad + c8 * 8;
                uint32_t lo = tb, hi = tb, seen_lo = 0, seen_hi = 0;
                if (PAIR)
                {
                    for (int32_t tt = 0; tt < I.t; ++tt)
                        first_nonzero(lo, seen_lo, *reinterpret_cast<const uint32_t *>(sp + tt * g.sh * g.sw * 4), tt);
                    *reinterpret_cast<uint32_t *>(fp) = lo;
                }
                else
                {
                    for (int32_t tt = 0; tt < I.t; ++tt)
                    {
                        const uint2 v = *reinterpret_cast<const uint2 *>(sp + tt * g.sh * g.sw * g.cin_pad);
                        first_nonzero(lo, seen_lo, v.x, tt);
                        first_nonzero(hi, seen_hi, v.y, tt);
                    }
                    *reinterpret_cast<uint2 *>(fp) = make_uint2(lo, hi);
                }
            }
        }

        // groups of nct region rows; wave (cgw, ct) takes row grp0 + ct
        for (int32_t grp0 = 0; grp0 < g.rh && oy0 + grp0 < I.ho; grp0 += g.nct)
        {
            const int32_t oy = oy0 + grp0 + ct, ox = ox0 + j;
            const bool pvalid = oy < I.ho && ox < I.wo;
            const uint8_t *wb = reinterpret_cast<const uint8_t *>(win16) + (ct * g.pitch + j) * UB;
            uint32_t mask[2][4];
            int32_t mxb[2][4];
            float fv[2][4];  // each value at its first firing step
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) mask[r][i] = 0, mxb[r][i] = 0, fv[r][i] = 0.0f;

            for (int32_t t0 = 0; t0 < I.t; t0 += g.tc)
            {
                const int32_t tc = min(g.tc, I.t - t0);
                // the FP16 windows of steps [t0, t0 + tc): stage rows
                // grp0 .. grp0 + wrows, columns xoff .. xoff + wcols (the
                // stage holds 0 wherever the input does not reach)
                if (wslot < wslots)
                {
                    // the slot walks the (step ts, row r) pairs by (wdq, wdr):
                    // its stage byte offset and window unit follow by additions
                    const int32_t rstride = g.sw * g.cin_pad, wstep = PAIR ? 2 * g.win_units : g.win_units;
                    int32_t r = wslot % g.wrows;
                    int32_t sof = ((grp0 + t0 * g.sh) * g.sw + xoff + wcc) * g.cin_pad + wc8 * 8 +
                                  ((wslot / g.wrows) * g.sh + r) * rstride;
                    int32_t wat = (wslot / g.wrows) * wstep + r * g.pitch + wcu;
                    const int32_t d_sof = (wdq * g.sh + wdr) * rstride, wrap_sof = (g.sh - g.wrows) * rstride;
                    const int32_t d_wat = wdq * wstep + wdr * g.pitch, wrap_wat = wstep - g.wrows * g.pitch;
                    const int32_t prs = tc * g.wrows;
                    // batches of WU (step, row) pairs a slot, their stage reads together
                    for (int32_t pr0 = wslot; pr0 < prs; pr0 += WU * wslots)
                    {
                        uint2 raw[WU];
                        int32_t at[WU];
#pragma unroll
                        for (int u = 0; u < WU; ++u)
                        {
                            raw[u] = make_uint2(0, 0);
                            if (pr0 + u * wslots < prs)
                            {
                                if (PAIR)
                                    raw[u].x = *reinterpret_cast<const uint32_t *>(stage + sof);
                                else
                                    raw[u] = *reinterpret_cast<const uint2 *>(stage + sof);
                            }
                            at[u] = wat;
                            r += wdr, sof += d_sof, wat += d_wat;
                            if (r >= g.wrows) r -= g.wrows, sof += wrap_sof, wat += wrap_wat;
                        }
#pragma unroll
                        for (int u = 0; u < WU; ++u)
                            if (pr0 + u * wslots < prs)
                            {
                                const uint4 hv = subnormals(raw[u]);
                                if (PAIR)
                                    reinterpret_cast<uint2 *>(win16)[at[u]] = make_uint2(hv.x, hv.y);
                                else
                                    win16[at[u]] = hv;
                            }
                    }
                }
                __syncthreads();
                // the chunk's steps without barriers
                // B operands: up to 8 K-steps two whole steps of buffers
                // (b0, b1), loaded two steps ahead of their use; longer chains
                // (a whole step of buffers spills at 13 K-steps) a ring of R
                // registers, operand s of a step in slot s % R, loaded as soon
                // as the slot's previous reader has issued, R MFMA pairs or
                // more ahead of its use
                constexpr bool RING = S > 8;
                constexpr int R = RING ? 4 : S;
                half8 b0[S], b1[RING ? 1 : S];
                float4v c0, c1, n0, n1;
                auto load1 = [&](half8 &dst, int s, const uint8_t *base) {
                    if (PAIR)
                    {
                        const uint2 lo = *reinterpret_cast<const uint2 *>(base + offs[s]);
                        const uint2 hi = *reinterpret_cast<const uint2 *>(base + offs[s] + 8);
                        dst = __builtin_bit_cast(half8, make_uint4(lo.x, lo.y, hi.x, hi.y));
                    }
                    else
                        dst = *reinterpret_cast<const half8 *>(base + offs[s]);
                };
                auto load = [&](half8(&b)[S], const uint8_t *base) {
#pragma unroll
                    for (int s = 0; s < S; ++s) load1(b[s], s, base);
                };
                // the ring's chain: the MFMAs of the step at `cur` into (u0,
                // u1), loading ahead (past the step's operands: the step at `next`)
                auto chain = [&](float4v &u0, float4v &u1, const uint8_t *cur, const uint8_t *next) {
                    u0 = float4v{0, 0, 0, 0}, u1 = float4v{0, 0, 0, 0};
#pragma unroll
                    for (int s = 0; s < S; ++s)
                    {
                        u0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0][s], b0[s % R], u0, 0, 0, 0);
                        u1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1][s], b0[s % R], u1, 0, 0, 0);
                        // slot s % R is next read by operand s + R of this
                        // step, or, past the step's last reader of the slot,
                        // by operand s % R of the next step
                        if (s + R < S)
                            load1(b0[s % R], s + R, cur);
                        else
                            load1(b0[s % R], s % R, next);
                    }
                };
                auto step_base = [&](int32_t k) { return wb + min(k, tc - 1) * win_step; };
                if constexpr (RING)
                {
#pragma unroll
                    for (int s = 0; s < R; ++s) load1(b0[s], s, wb);
                    chain(c0, c1, step_base(0), step_base(1));
                }
                else
                {
                    load(b0, wb);
                    load(b1, step_base(1));
                }
                auto update = [&](int32_t k, const float4v &u0, const float4v &u1) {
                    step_update_first(u0, u1, g.th, mask, mxb, fv);
                    const int64_t st = t0 + k;
                    if (zspk && oy0 + grp0 + zr < I.ho)
                    {
                        uint8_t *d = out_i + ((st * g.cout + zc) * I.ho + oy0 + grp0 + zr) * I.wo + ox0;
                        if (spk16)
                            *reinterpret_cast<u32x4 *>(d) = u32x4{0u, 0u, 0u, 0u};
                        else
                            for (int32_t x = 0; x < RW && ox0 + x < I.wo; ++x) d[x] = 0;
                    }
                    if (zpool && (oy0 + grp0) / ro.pkh + zr < I.pho)
                    {
                        const int32_t px0 = ox0 / ro.pkw;
                        uint8_t *d = pool_i + ((st * g.cout + zc) * I.pho + (oy0 + grp0) / ro.pkh + zr) * I.pwo + px0;
                        if (pool8 && px0 + 8 <= I.pwo)
                            *reinterpret_cast<uint64_t *>(d) = 0;
                        else
                            for (int32_t x = 0; x < pbytes && px0 + x < I.pwo; ++x) d[x] = 0;
                    }
                    if (DIAG && pvalid)
#pragma unroll
                        for (int r = 0; r < 2; ++r)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                            {
                                const int32_t o = o0 + r * 16 + i;
                                if (o < g.cout)
                                    pot_out[((int64_t(itm.x) * I.t + t0 + k) * g.cout + o) * (I.ho * I.wo) +
                                            oy * I.wo + ox] = (r == 0 ? u0[i] : u1[i]) * TWO_POW_24;
                            }
                };
                if (!RING)
                    // two steps an iteration: both steps' MFMAs, each chain's
                    // operands reloaded for two steps later as soon as it has
                    // issued, then both updates (no accumulator lives across
                    // iterations; loads past the chunk re-read its last step)
                    for (int32_t k = 0; k < tc; k += 2)
                    {
                        if constexpr (!RING)
                        {
                            mfma_chain(a, b0, c0, c1);
                            load(b0, step_base(k + 2));
                            if (k + 1 < tc) mfma_chain(a, b1, n0, n1);
                            load(b1, step_base(k + 3));
                            update(k, c0, c1);
                            if (k + 1 < tc) update(k + 1, n0, n1);
                        }
                    }
                else
                    // long chains: one accumulator pair (the other wave of the
                    // SIMD covers the wait for the last MFMA's result)
                    for (int32_t k = 0; k < tc; ++k)
                    {
                        update(k, c0, c1);
                        if (k + 1 < tc) chain(c0, c1, step_base(k + 1), step_base(k + 2));
                    }
                if (t0 + tc < I.t) __syncthreads();  // windows free for the next chunk
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
            // the lane's B operands for pixel (oy, ox) at step st, from the
            // stage: the bytes the step loop's window held for that step
            auto gather1 = [&](int s, int32_t st) {
                const int4 te = gtab[4 * s + q];
                const int32_t iy = oy + te.x, ix = ox + te.y;
                uint2 raw = make_uint2(0, 0);
                const bool row = st >= 0 && iy >= 0 && iy < I.h;
                const uint8_t *sp = stage + ((st * g.sh + iy - sy0) * g.sw + ix - sx0) * g.cin_pad + te.z;
                if (PAIR)
                {
                    // the two taps' 4 channels, each 0 outside the input
                    if (row && ix >= 0 && ix < I.w) raw.x = *reinterpret_cast<const uint32_t *>(sp);
                    if (row && ix + 1 >= 0 && ix + 1 < I.w) raw.y = *reinterpret_cast<const uint32_t *>(sp + 4);
                }
                else if (row && ix >= 0 && ix < I.w)
                    raw = *reinterpret_cast<const uint2 *>(sp);
                return __builtin_bit_cast(half8, subnormals(raw));
            };
            // the step loop's MFMA sequence on the gathered operands at step
            // st, gathered GR K-steps ahead of their use
            auto gchain = [&](int32_t st, float4v &u0, float4v &u1) {
                constexpr int GR = S < 4 ? S : 4;
                half8 gb[GR];
#pragma unroll
                for (int s = 0; s < GR; ++s) gb[s] = gather1(s, st);
                u0 = float4v{0, 0, 0, 0}, u1 = float4v{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < S; ++s)
                {
                    u0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0][s], gb[s % GR], u0, 0, 0, 0);
                    u1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1][s], gb[s % GR], u1, 0, 0, 0);
                    if (s + GR < S) gb[s % GR] = gather1(s + GR, st);
                }
            };
            // inhibit's kept channel c* at its step ts = T - max(cnt, 1).  When
            // the pixel's fired steps (all channels) form one run up to the
            // last step, ts is the first step any channel fired: a value that
            // fired at ts fired first there (fv holds it) and any other is 0
            // after the threshold.  Otherwise the step loop's MFMA sequence on
            // the gathered operands at ts gives the values, bit for bit.
            int32_t ts;
            bool run;
            {
                uint32_t all = 0;
                for (int32_t k = 0; k < g.cg; ++k) all |= any_s[(ct * g.cg + k) * COLS + j];
                const int32_t cnt = __builtin_popcount(all);
                ts = I.t - (cnt > 1 ? cnt : 1);
                run = (all & (all + 1u)) == 0u;
            }
            // pass 1 where needed: the potentials at inhibit's step; the kept
            // channel is the first maximum of the thresholded values
            // (inhibit.cpp:64-85)
            {
                float vt[2][4];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) vt[r][i] = fv[r][i];
                if (__builtin_amdgcn_ballot_w64(!run))
                {
                    float4v g0, g1;
                    gchain(ts, g0, g1);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (!run) vt[0][i] = g0[i], vt[1][i] = g1[i];
                }
                const int32_t sh = I.t - 1 - ts;
                float bv = -1.0f;
                int32_t bi = INT32_MAX;
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                    {
                        const float xp = (mask[r][i] >> sh) & 1u ? vt[r][i] : 0.0f;
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
                float m = red_v[ct * g.cg * COLS + j];
                cs = red_i[ct * g.cg * COLS + j];
                for (int32_t k = 1; k < g.cg; ++k)
                    combine(m, cs, red_v[(ct * g.cg + k) * COLS + j], red_i[(ct * g.cg + k) * COLS + j]);
            }
            // the lane holding (c*, pixel) has c*'s fire mask, maximum and
            // value at its first firing step (convwta's t0 value)
            const bool own = (cs >> 5) == cgw && ((cs >> 2) & 3) == q;
            uint32_t om = 0;
            int32_t oxb = 0;
            float first = 0.0f;
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (((cs >> 4) & 1) == r && (cs & 3) == i) om = mask[r][i], oxb = mxb[r][i], first = fv[r][i];
            const float pmax = om ? __builtin_bit_cast(float, oxb) * TWO_POW_24 : 0.0f;
            if (own)
            {
                const uint32_t fwd = om ? __builtin_bitreverse32(om) >> (32 - I.t) : 0u;
                if (pvalid) summary[I.sum_off + int64_t(oy) * I.wo + ox] = Summary{cs, fwd, om ? first * TWO_POW_24 : 0.0f, pmax};
                sel_c[ct * COLS + j] = cs, sel_m[ct * COLS + j] = pvalid ? fwd : 0u;
            }
            {
                const float mv = wave_max(own && pvalid && pmax > 0.0f ? pmax : 0.0f);
                if (itm.x != imx_inst) flush_imax(), imx_inst = itm.x, imx = 0.0f;
                imx = imx < mv ? mv : imx;
            }
            if (zs + zp > 0)
            {
                // the group's 1s: pixel (row, column) at step t in its kept
                // channel's plane and pooled window, over the zeros (the
                // barrier orders every zero store of the block before them)
                __syncthreads();
                const int32_t per_row = COLS * I.t;
                for (int32_t e = threadIdx.x; e < g.nct * per_row; e += blockDim.x)
                {
                    const int32_t r = e / per_row, px = (e % per_row) / I.t, st = (e % per_row) % I.t;
                    if (!((sel_m[r * COLS + px] >> st) & 1u)) continue;
                    const int32_t c = sel_c[r * COLS + px], py = oy0 + grp0 + r, pxx = ox0 + px;
                    if (out_i) out_i[((int64_t(st) * g.cout + c) * I.ho + py) * I.wo + pxx] = 1;
                    if (pool_i && py / ro.pkh < I.pho && pxx / ro.pkw < I.pwo)
                        pool_i[((int64_t(st) * g.cout + c) * I.pho + py / ro.pkh) * I.pwo + pxx / ro.pkw] = 1;
                }
            }
        }
    }
    flush_imax();
}

// convwta over one instance's summaries (fused_select's rule, wta.cpp:49-64,
// 128-150) for planes of any size, one block per instance.  The candidate of
// pixel p is (v, c * hw + p) with v = (T-1-t0) + first/diff (-FLT_MAX when it
// never fired), diff = 2 * max(0, the instance's potentials) and c its kept
// channel; a round's winner is the first maximum
// over the candidates not excluded by an earlier winner (same channel, or
// inside its box).  Each thread keeps the SEL_TOP best candidates of its
// pixels in registers, ordered, from one pass; a round takes every thread's
// first candidate that no winner excludes.  A candidate outside the list is
// excluded or ranks below all of the list, so that is the thread's best
// unless every listed candidate is excluded; then (if the thread has more
// pixels than the list) it refills the list with its best SEL_TOP candidates
// that no winner excludes.  Winners only accumulate, so the rule holds for
// the refilled list in later rounds.  wkey[slot] = channel << 16 | pixel, -1
// past the instance's winners.
constexpr int SEL_TOP = 8;
constexpr int SEL_MAX = 256;  // winners an instance can have on the region path
constexpr int SEL_BATCH = 8;  // summaries a thread loads together

__global__ void __launch_bounds__(1024)
    region_select(const Summary *__restrict__ summary, const RegionInst *__restrict__ inst,
                  const uint32_t *__restrict__ imax, int32_t ry, int32_t rx, int32_t count,
                  spk_winner *__restrict__ win, int64_t *__restrict__ counts, int32_t *__restrict__ wkey)
{
    __shared__ ArgMax am[16];
    __shared__ int32_t wz[SEL_MAX], wy[SEL_MAX], wx[SEL_MAX];  // the winners so far
    const int64_t s = blockIdx.x;
    const RegionInst I = inst[s];
    const int32_t hw = I.ho * I.wo;
    const Summary *sm = summary + I.sum_off;
    const int32_t bd = int32_t(blockDim.x);
    // the instance's maximum potential (from 0, as the reference's scan),
    // which the region kernel recorded
    const float m = __builtin_bit_cast(float, imax[s]);
    const float diff = (m - 0.0f) * 2.0f;
    auto value = [&](const Summary &e, int32_t p, float &v, int32_t &key) {
        v = e.mask ? float(I.t - int32_t(__builtin_ctz(e.mask)) - 1) + (e.first - 0.0f) / diff : -FLT_MAX;
        key = e.c * hw + p;
    };
    auto excluded = [&](int32_t key, int32_t k) {
        const int32_t z = key / hw, p = key % hw, py = p / I.wo, px = p % I.wo;
        for (int32_t r = 0; r < k; ++r)
            if (wz[r] == z || (abs(py - wy[r]) <= ry && abs(px - wx[r]) <= rx)) return true;
        return false;
    };
    float tv[SEL_TOP];
    int32_t tk[SEL_TOP];
    const int32_t mine = hw > int32_t(threadIdx.x) ? (hw - int32_t(threadIdx.x) + bd - 1) / bd : 0;  // its pixels
    // the thread's SEL_TOP best candidates that none of the first k winners
    // excludes, ordered
    auto fill = [&](int32_t k) {
#pragma unroll
        for (int u = 0; u < SEL_TOP; ++u) tv[u] = -INFINITY, tk[u] = INT32_MAX;
        for (int32_t q0 = 0; q0 < mine; q0 += SEL_BATCH)
        {
            Summary e[SEL_BATCH];
#pragma unroll
            for (int u = 0; u < SEL_BATCH; ++u)
                e[u] = q0 + u < mine ? sm[int32_t(threadIdx.x) + (q0 + u) * bd] : Summary{0, 0, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < SEL_BATCH; ++u)
            {
                if (q0 + u >= mine) break;
                float v;
                int32_t key;
                value(e[u], int32_t(threadIdx.x) + (q0 + u) * bd, v, key);
                if (!better(v, key, tv[SEL_TOP - 1], tk[SEL_TOP - 1]) || (k > 0 && excluded(key, k))) continue;
                // insertion into the ordered list
#pragma unroll
                for (int w = SEL_TOP - 1; w > 0; --w)
                {
                    const bool shift = better(v, key, tv[w - 1], tk[w - 1]);
                    const bool here = !shift && better(v, key, tv[w], tk[w]);
                    const float nv = shift ? tv[w - 1] : (here ? v : tv[w]);
                    const int32_t nk = shift ? tk[w - 1] : (here ? key : tk[w]);
                    tv[w] = nv, tk[w] = nk;
                }
                if (better(v, key, tv[0], tk[0])) tv[0] = v, tk[0] = key;
            }
        }
    };
    fill(0);
    int32_t k = 0;
    for (; k < count; ++k)
    {
        ArgMax a{-INFINITY, INT32_MAX};
        bool found = false;
#pragma unroll
        for (int u = 0; u < SEL_TOP; ++u)
            if (!found && tk[u] != INT32_MAX && !excluded(tk[u], k)) a.v = tv[u], a.i = tk[u], found = true;
        if (!found && mine > SEL_TOP)
        {
            fill(k);
            if (tk[0] != INT32_MAX) a.v = tv[0], a.i = tk[0];
        }
        a = block_argmax(a, am);
        if (fabsf(a.v - (-FLT_MAX)) < 1e-5f || a.i == INT32_MAX) break;
        const int32_t z = a.i / hw, y = (a.i % hw) / I.wo, x = (a.i % hw) % I.wo;
        if (threadIdx.x == 0)
        {
            win[s * count + k] = spk_winner{0, 0, z, y, x}, wkey[s * count + k] = z << 16 | (a.i % hw);
            wz[k] = z, wy[k] = y, wx[k] = x;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[s] = k;
    for (int32_t r = k + int32_t(threadIdx.x); r < count; r += blockDim.x)
        wkey[s * count + r] = -1, win[s * count + r] = spk_winner{0, 0, 0, 0, 0};
}

// Writes channel planes 0 .. cout-1 (at most 64) of 4 consecutive positions:
// bit ch of cm[p] is position p's byte in plane ch.  Byte g of the four masks
// side by side holds channels 8g .. 8g+7 of the 4 positions, so plane 8g + i
// is that dword's bit i of every byte.
__device__ __forceinline__ void store_planes(uint8_t *d, int64_t plane, int32_t cout, const uint64_t (&cm)[4])
{
    for (int32_t g = 0; 8 * g < cout; ++g)
    {
        const uint32_t x = uint32_t((cm[0] >> (8 * g)) & 0xFF) | uint32_t((cm[1] >> (8 * g)) & 0xFF) << 8 |
                           uint32_t((cm[2] >> (8 * g)) & 0xFF) << 16 | uint32_t((cm[3] >> (8 * g)) & 0xFF) << 24;
        const int32_t n = min(8, cout - 8 * g);
        for (int32_t i = 0; i < n; ++i)
            __builtin_nontemporal_store((x >> i) & 0x01010101u, reinterpret_cast<uint32_t *>(d + (8 * g + i) * plane));
    }
}

// Spikes [t][cout][ho][wo] of every instance from the summaries, when the
// region kernel does not write them.  Grid
// (chunk, instance): a thread owns 4 consecutive pixels (one when hw % 4 !=
// 0 or the instance's spikes are not 4-aligned), reads their summaries once
// and writes every step's channel planes, one 32-bit store per plane.  The
// 32-bit stores are non-temporal: nothing here reads the output again.
constexpr int WRITE_THREADS = 256;

__global__ void __launch_bounds__(WRITE_THREADS)
    region_spikes(const Summary *__restrict__ summary, const RegionInst *__restrict__ inst, int32_t cout,
                  uint8_t *__restrict__ out)
{
    const RegionInst I = inst[blockIdx.y];
    const int32_t hw = I.ho * I.wo;
    uint8_t *base = out + I.out_off;
    const bool quad = hw % 4 == 0 && (reinterpret_cast<uintptr_t>(base) & 3) == 0;
    const int32_t per = quad ? 4 : 1, units = hw / per;
    const int32_t u = int32_t(blockIdx.x) * WRITE_THREADS + int32_t(threadIdx.x);
    if (u >= units) return;
    const Summary *sm = summary + I.sum_off + int64_t(u) * per;
    int32_t c[4];
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const Summary e = k < per ? sm[k] : Summary{-1, 0, 0.0f, 0.0f};
        c[k] = e.c, m[k] = e.mask;
    }
    const int64_t plane = hw;
    for (int32_t step = 0; step < I.t; ++step)
    {
        int32_t a[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = (m[k] >> step) & 1u ? c[k] : -1;
        uint8_t *d = base + int64_t(step) * cout * plane + int64_t(u) * per;
        if (quad && cout <= 64)
        {
            uint64_t cm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) cm[k] = a[k] >= 0 ? uint64_t(1) << a[k] : 0;
            store_planes(d, plane, cout, cm);
        }
        else if (quad)
            for (int32_t ch = 0; ch < cout; ++ch)
                __builtin_nontemporal_store(uint32_t(a[0] == ch) | uint32_t(a[1] == ch) << 8 |
                                                uint32_t(a[2] == ch) << 16 | uint32_t(a[3] == ch) << 24,
                                            reinterpret_cast<uint32_t *>(d + ch * plane));
        else
            for (int32_t ch = 0; ch < cout; ++ch) d[ch * plane] = a[0] == ch;
    }
}

// Pooled spikes [t][cout][pho][pwo].  With windows of at most 4 pixels, at
// most 64 channels and 4-aligned planes, a thread owns 4 consecutive pooled
// positions: their windows' summaries stay in registers and each step builds
// one 64-bit channel mask per position (a pixel has at most one active
// channel), then writes one 32-bit store per channel plane.  Otherwise a
// thread owns one position and compares channel by channel.
__global__ void __launch_bounds__(WRITE_THREADS)
    region_pooled(const Summary *__restrict__ summary, const RegionInst *__restrict__ inst, int32_t cout,
                  uint8_t *__restrict__ pooled, int32_t pkh, int32_t pkw, int32_t psh, int32_t psw)
{
    const RegionInst I = inst[blockIdx.y];
    const int32_t phw = I.pho * I.pwo, win = pkh * pkw;
    uint8_t *base = pooled + I.pool_off;
    const Summary *sm = summary + I.sum_off;
    const bool quad = win <= 4 && cout <= 64 && phw % 4 == 0 && (reinterpret_cast<uintptr_t>(base) & 3) == 0;
    const int32_t u = int32_t(blockIdx.x) * WRITE_THREADS + int32_t(threadIdx.x);
    auto load = [&](int32_t pos, int32_t k, int32_t &c, uint32_t &m) {
        const int32_t py = pos / I.pwo, px = pos % I.pwo, y = py * psh + k / pkw, x = px * psw + k % pkw;
        const bool in = k < win && y < I.ho && x < I.wo;
        const Summary e = in ? sm[y * I.wo + x] : Summary{-1, 0, 0.0f, 0.0f};
        c = e.c, m = e.mask;
    };
    if (quad)
    {
        if (u >= phw / 4) return;
        int32_t wc[4][4];
        uint32_t wm[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int k = 0; k < 4; ++k) load(4 * u + p, k, wc[p][k], wm[p][k]);
        uint8_t *d = base + 4 * int64_t(u);
        for (int32_t step = 0; step < I.t; ++step, d += int64_t(cout) * phw)
        {
            uint64_t cm[4];
#pragma unroll
            for (int p = 0; p < 4; ++p)
            {
                cm[p] = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    cm[p] |= uint64_t((wm[p][k] >> step) & 1u) << (wc[p][k] & 63);
            }
            store_planes(d, phw, cout, cm);
        }
        return;
    }
    if (u >= phw) return;
    int32_t wc[16];
    uint32_t wm[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) load(u, k, wc[k], wm[k]);
    uint8_t *d = base + u;
    for (int32_t step = 0; step < I.t; ++step, d += int64_t(cout) * phw)
    {
        int32_t a[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = (wm[k] >> step) & 1u ? wc[k] : -1;
        for (int32_t ch = 0; ch < cout; ++ch)
        {
            bool hit = false;
#pragma unroll
            for (int k = 0; k < 16; ++k) hit |= a[k] == ch;
            d[int64_t(ch) * phw] = hit;
        }
    }
}

// Potentiate/depress bits of every winner slot (fused_records with the
// instance tables): bit i of slot e is isum <= osum for weight i = (ch, ky,
// kx) of the winner's kernel (stdp.cpp:88-118).
__global__ void __launch_bounds__(256)
    region_records(const uint8_t *__restrict__ fs, const Summary *__restrict__ summary,
                   const RegionInst *__restrict__ inst, const int32_t *__restrict__ wkey, FusedStdp p, int32_t pt,
                   int32_t pl, uint32_t *__restrict__ bits)
{
    const int64_t e = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (e >= p.slots) return;
    const int32_t k = wkey[e];
    if (k < 0) return;
    const int lane = threadIdx.x & 63;
    const RegionInst I = inst[e / p.count];
    const int32_t pos = k & 0xFFFF, y = pos / I.wo - pt, x = pos % I.wo - pl;
    const uint32_t mask = summary[I.sum_off + pos].mask;
    const int32_t osum = mask ? int32_t(__builtin_ctz(mask)) : I.t;
    const int32_t taps = p.kh * p.kw, ksz = p.cin * taps;
    const uint8_t *fsn = fs + I.fs_off;
    uint32_t *rec = bits + e * p.words;
    for (int32_t base = 0; base < ksz; base += 64)
    {
        const int32_t i = base + lane;
        bool bit = false;
        if (i < ksz)
        {
            const int32_t ch = i / taps, iy = y + (i / p.kw) % p.kh, ix = x + i % p.kw;
            const int32_t isum =
                (iy >= 0 && iy < I.h && ix >= 0 && ix < I.w) ? fsn[(iy * I.w + ix) * p.cin_pad + ch] : I.t;
            bit = isum <= osum;
        }
        const uint64_t b = __ballot(bit);
        if (lane == 0 && base / 32 < p.words) rec[base / 32] = uint32_t(b);
        if (lane == 32 && base / 32 + 1 < p.words) rec[base / 32 + 1] = uint32_t(b >> 32);
    }
}

__global__ void __launch_bounds__(256) zero_words(uint32_t *__restrict__ p, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// A candidate launch configuration of the region kernel.
struct RegionConfig
{
    RegionGeom g;
    size_t lds = 0;
    int grid = 0;  // blocks: one a compute unit, at most one an item
    int2 *items = nullptr;  // device table [g.items] of (instance, region)
    // the instances in PIPE_CHUNKS chunks of about equal region counts: chunk
    // c is items [cut_item[c], cut_item[c+1]), instances [cut_inst[c], cut_inst[c+1])
    std::vector<int64_t> cut_item, cut_inst;
};

// Chunks of instances a step's region kernel runs in when the separate
// writers follow it: each chunk's writers overlap the next chunk's kernel.
constexpr int PIPE_CHUNKS = 8;

template <int S, bool DIAG, bool PAIR, bool BSTAGE>
void launch_region(const uint8_t *in, const half8 *frag, const RegionInst *inst, const RegionConfig &rc,
                   Summary *summary, float *pot, uint8_t *fs, const RegionOut &ro, int grid, hipStream_t s)
{
    static bool configured = false;  // per instantiation; the attribute is per function
    if (!configured)
    {
        SPK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_region<S, DIAG, PAIR, BSTAGE>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES));
        configured = true;
    }
    fused_region<S, DIAG, PAIR, BSTAGE><<<dim3(grid), dim3(64 * rc.g.waves), rc.lds, s>>>(in, frag, inst, rc.items,
                                                                                        summary, pot, fs, ro, rc.g);
    SPK_LAUNCHED();
}

// bstage: the buffer-load staging walk, denser than the flat one (faster
// alone, but it slows writer kernels running beside the kernel); the
// potentials variant always stages by flat loads.
void dispatch_region(const uint8_t *in, const half8 *frag, const RegionInst *inst, const RegionConfig &rc,
                     Summary *summary, float *pot, uint8_t *fs, const RegionOut &ro, int grid, hipStream_t s,
                     bool bstage)
{
    if (rc.g.pair)
        switch (rc.g.steps)
        {
#define CASE(k)                                                                              \
    case k:                                                                                  \
        return pot      ? launch_region<k, true, true, false>(in, frag, inst, rc, summary, pot, fs, ro, grid, s) \
               : bstage ? launch_region<k, false, true, true>(in, frag, inst, rc, summary, pot, fs, ro, grid, s) \
                        : launch_region<k, false, true, false>(in, frag, inst, rc, summary, pot, fs, ro, grid, s);
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default: SPK_ARG(false, "Fused layer: K-steps out of range.");
        }
    switch (rc.g.steps)
    {
#define CASE(k)                                                                               \
    case k:                                                                                   \
        return pot      ? launch_region<k, true, false, false>(in, frag, inst, rc, summary, pot, fs, ro, grid, s) \
               : bstage ? launch_region<k, false, false, true>(in, frag, inst, rc, summary, pot, fs, ro, grid, s) \
                        : launch_region<k, false, false, false>(in, frag, inst, rc, summary, pot, fs, ro, grid, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: SPK_ARG(false, "Fused layer: K-steps out of range.");
    }
}

template <typename T>
struct DeviceBuffer
{
    T *p = nullptr;
    explicit DeviceBuffer(int64_t count = 0) { reset(count); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
    void reset(int64_t count)
    {
        if (p) (void)hipFree(p), p = nullptr;
        if (count > 0) SPK_HIP(hipMalloc(reinterpret_cast<void **>(&p), size_t(count) * sizeof(T)));
    }
};
}  // namespace

// A fixed-shape layer step: instance table, candidate configurations and
// every work buffer, made once.
struct LayerPlan
{
    int device = 0;
    int64_t n = 0;
    int32_t cin = 0, cout = 0, kh = 0, kw = 0, pt = 0, pl = 0, tmax = 0, count = 0, ry = 0, rx = 0;
    int32_t pkh = 0, pkw = 0, psh = 0, psw = 0;
    bool pool = false, tune = false, tuned = false, fast = false;
    float th = 0;
    std::vector<RegionInst> host;
    int64_t in_bytes = 0, out_bytes = 0, pool_bytes = 0, pixels = 0, fs_bytes = 0, max_hw = 0, max_phw = 0;
    std::vector<std::unique_ptr<RegionConfig>> configs;
    std::vector<std::unique_ptr<DeviceBuffer<int2>>> item_tables;
    int chosen = 0;
    bool fuse_out = true;  // the region kernel writes the outputs (else the separate writers)
    DeviceBuffer<RegionInst> inst;
    DeviceBuffer<half8> frag;
    DeviceBuffer<Summary> summary;
    DeviceBuffer<uint8_t> fs;
    DeviceBuffer<int32_t> wkey;
    DeviceBuffer<uint32_t> imax;  // [n] the instances' maximum potentials (RegionOut::imax)
    DeviceBuffer<uint32_t> bits;
    FusedStdp stdp{};

    void run(const uint8_t *in, float *kernel, const spk_stdp_config *cfg, uint8_t *out, uint8_t *pooled,
             spk_winner *winners, int64_t *counts, float *pot, hipStream_t s);
    void choose(const uint8_t *in, uint8_t *out, uint8_t *pooled, hipStream_t s);
    void write_step(const uint8_t *in, const RegionConfig &rc, bool fused, uint8_t *out, uint8_t *pooled, float *pot,
                    uint8_t *fsp, hipStream_t s, SideStream *side);
    void launch_select(spk_winner *winners, int64_t *counts, hipStream_t s, int64_t i0, int64_t i1);
    void launch_writers(bool fused, const RegionConfig &rc, uint8_t *out, uint8_t *pooled, hipStream_t ws, int64_t i0,
                        int64_t i1);
    bool pool_in_kernel(const RegionConfig &rc, bool fused, const uint8_t *pooled) const;
};

namespace
{
// Geometry of one candidate: region rows = nct * m.
bool region_config(const LayerPlan &P, int32_t m, bool fast, RegionGeom &g, size_t &lds)
{
    g = RegionGeom{};
    g.fast = fast;
    g.cin = P.cin, g.cout = P.cout, g.kh = P.kh, g.kw = P.kw, g.pt = P.pt, g.pl = P.pl;
    const int32_t np = (P.kw + 1) / 2;
    g.pair = use_pair(P.cin, P.kh, P.kw);
    g.cin_pad = g.pair ? 4 : int32_t(cdiv(P.cin, 8) * 8);
    const int32_t planes = g.pair ? 1 : g.cin_pad / 8;
    g.groups = g.pair ? P.kh * np : P.kh * P.kw * planes;
    g.steps = int32_t(cdiv(g.groups, 4));
    const int32_t kspan = g.pair ? 2 * np : P.kw;  // window columns a pixel's taps span
    g.cg = int32_t(cdiv(P.cout, 32));
    g.nct = std::max(1, REGION_WAVES / g.cg);
    g.waves = g.cg * g.nct;
    g.rh = g.nct * m;
    g.sh = g.rh + P.kh - 1;
    g.sw = int32_t(cdiv(RW + kspan - 1 + 3, 4) * 4);
    g.wrows = g.nct + P.kh - 1;
    g.wcols = RW + kspan - 1;
    g.pitch = planes * g.wcols;
    g.win_units = g.pair ? int32_t(cdiv(g.wrows * g.pitch, 2)) : g.wrows * g.pitch;
    g.th = P.th * (1.0f / TWO_POW_24);  // potentials are computed times 2^-24
    const int64_t stage = cdiv(int64_t(P.tmax) * g.sh * g.sw * g.cin_pad, 16) * 16;
    const int64_t fixed = stage + region_fixed(g.waves, g.nct, g.steps);
    const int64_t room = (LDS_BYTES - fixed) / (int64_t(g.win_units) * 16);
    if (room < 1) return false;
    // thread roles: a window row's units, and one stage row's dwords (or bytes)
    if (g.pitch > 64 * g.waves) return false;
    if (fast && g.sw * g.cin_pad / 4 > 64 * g.waves) return false;
    const int64_t chunks = cdiv(P.tmax, std::min<int64_t>(room, P.tmax));
    g.tc = int32_t(cdiv(P.tmax, chunks));
    g.stage_bytes = int32_t(stage);
    lds = size_t(fixed + int64_t(g.tc) * g.win_units * 16);
    return true;
}
}  // namespace

void LayerPlan::choose(const uint8_t *in, uint8_t *out, uint8_t *pooled, hipStream_t s)
{
    // time every candidate with the outputs written inside the region kernel
    // and with the separate writers, on this input, and keep the fastest (the
    // run that follows rewrites the summaries, map and outputs)
    float best = INFINITY;
    hipEvent_t e0, e1;
    SPK_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
    SPK_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
    for (int fused = 1; fused >= 0; --fused)
        for (size_t c = 0; c < configs.size(); ++c)
        {
            float ms = 0.0f;
            for (int rep = 0; rep < 2; ++rep)  // the first launch configures and warms
            {
                SPK_HIP(hipEventRecord(e0, s));
                {
                    SideStream side(s, fused == 0);
                    write_step(in, *configs[c], fused != 0, out, pooled, nullptr, fs.p, s, &side);
                    side.join();
                }
                SPK_HIP(hipEventRecord(e1, s));
                SPK_HIP(hipEventSynchronize(e1));
                SPK_HIP(hipEventElapsedTime(&ms, e0, e1));
            }
            if (ms < best) best = ms, chosen = int(c), fuse_out = fused != 0;
        }
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    tuned = true;
}

// The region kernel with config rc and the step's outputs: written inside the
// kernel (the pool only when its windows lie inside one group of region
// rows), or by the separate writers on
// the side stream (on `s` after the kernel without one).  With the separate
// writers the kernel runs over the instances in chunks, and each chunk's
// writers start as soon as its summaries are there, overlapping the next
// chunk's kernel (which moves little HBM while the writers move little
// else); the caller joins the side stream.  Only a kernel with no writers
// beside it stages by buffer loads.
bool LayerPlan::pool_in_kernel(const RegionConfig &rc, bool fused, const uint8_t *pooled) const
{
    return fused && pool && pooled && psh == pkh && psw == pkw && rc.g.nct % pkh == 0 && RW % pkw == 0;
}

void LayerPlan::write_step(const uint8_t *in, const RegionConfig &rc, bool fused, uint8_t *out, uint8_t *pooled,
                           float *pot, uint8_t *fsp, hipStream_t s, SideStream *side)
{
    const RegionOut ro{fused ? out : nullptr, pool_in_kernel(rc, fused, pooled) ? pooled : nullptr, pkh, pkw, imax.p};
    const bool writers = (out && !fused) || (pool && pooled && !pool_in_kernel(rc, fused, pooled));
    const int64_t chunks = int64_t(rc.cut_item.size()) - 1;
    if (!writers || !side || pot || chunks < 2)
    {
        dispatch_region(in, frag.p, inst.p, rc, summary.p, pot, fsp, ro, rc.grid, s, true);
        if (writers)
        {
            if (side) side->again();
            launch_writers(fused, rc, out, pooled, side ? side->stream : s, 0, n);
        }
        return;
    }
    // a chunk's writers on the side stream, beside the next chunks' kernels
    for (int64_t c = 0; c < chunks; ++c)
    {
        RegionConfig part;
        part.g = rc.g, part.lds = rc.lds, part.items = rc.items + rc.cut_item[size_t(c)];
        part.g.items = rc.cut_item[size_t(c) + 1] - rc.cut_item[size_t(c)];
        dispatch_region(in, frag.p, inst.p, part, summary.p, nullptr, fsp, ro,
                        int(std::min<int64_t>(rc.grid, part.g.items)), s, false);
        side->again();
        launch_writers(fused, rc, out, pooled, side->stream, rc.cut_inst[size_t(c)], rc.cut_inst[size_t(c) + 1]);
    }
}

// convwta of instances [i0, i1).
void LayerPlan::launch_select(spk_winner *winners, int64_t *counts, hipStream_t s, int64_t i0, int64_t i1)
{
    if (i1 <= i0) return;
    region_select<<<dim3(unsigned(i1 - i0)), 1024, 0, s>>>(summary.p, inst.p + i0, imax.p + i0, ry, rx, count,
                                                           winners + i0 * count, counts + i0, wkey.p + i0 * count);
    SPK_LAUNCHED();
}

// The separate writers for instances [i0, i1).
void LayerPlan::launch_writers(bool fused, const RegionConfig &rc, uint8_t *out, uint8_t *pooled, hipStream_t ws,
                               int64_t i0, int64_t i1)
{
    const bool pool_in = pool_in_kernel(rc, fused, pooled);
    if (i1 <= i0) return;
    if (out && !fused)
    {
        region_spikes<<<dim3(unsigned(cdiv(max_hw, WRITE_THREADS)), unsigned(i1 - i0)), WRITE_THREADS, 0, ws>>>(
            summary.p, inst.p + i0, cout, out);
        SPK_LAUNCHED();
    }
    if (pool && pooled && !pool_in)
    {
        region_pooled<<<dim3(unsigned(cdiv(max_phw, WRITE_THREADS)), unsigned(i1 - i0)), WRITE_THREADS, 0, ws>>>(
            summary.p, inst.p + i0, cout, pooled, pkh, pkw, psh, psw);
        SPK_LAUNCHED();
    }
}

void LayerPlan::run(const uint8_t *in, float *kernel, const spk_stdp_config *cfg, uint8_t *out, uint8_t *pooled,
                    spk_winner *winners, int64_t *counts, float *pot, hipStream_t s)
{
    const RegionGeom &g0 = configs[0]->g;
    const dim3 fgrid(unsigned(cdiv(int64_t(g0.cg) * 2 * g0.steps * 64, 256)));
    if (g0.pair)
        prep_fragments_pair<<<fgrid, 256, 0, s>>>(kernel, frag.p, g0.cout, g0.cin, g0.kh, g0.kw, g0.groups, g0.steps,
                                                  g0.cg * 2);
    else
        prep_fragments<<<fgrid, 256, 0, s>>>(kernel, frag.p, g0.cout, g0.cin, g0.kh, g0.kw, g0.cin_pad, g0.groups,
                                             g0.steps, g0.cg * 2, 0);
    SPK_LAUNCHED();
    if (tune && !tuned) choose(in, out, pooled, s);
    const bool learn = cfg != nullptr;
    KernelTimer &tm = timer();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tm.on)
    {
        SPK_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
        SPK_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        SPK_HIP(hipEventRecord(e0, s));
    }
    const RegionConfig &rc = *configs[chosen];
    // outputs the kernel does not write: the separate writers on the side
    // stream, beside the kernel's later chunks
    const bool sep = (out && !fuse_out) || (pool && pooled && !pool_in_kernel(rc, fuse_out, pooled));
    zero_words<<<dim3(unsigned(cdiv(n, 256))), 256, 0, s>>>(imax.p, n);  // a kernel: capturable
    SPK_LAUNCHED();
    SideStream side(s, sep);
    write_step(in, rc, fuse_out, out, pooled, pot, learn ? fs.p : nullptr, s, sep ? &side : nullptr);
    if (tm.on)
    {
        SPK_HIP(hipEventRecord(e1, s));
        tm.events.emplace_back(e0, e1);
    }
    // convwta on the kernel's stream, beside the last chunk's writers
    launch_select(winners, counts, s, 0, n);
    side.join();  // the winners, for the records
    if (learn)
    {
        region_records<<<dim3(unsigned(cdiv(stdp.slots, 4))), 256, 0, s>>>(fs.p, summary.p, inst.p, wkey.p, stdp,
                                                                          pt, pl, bits.p);
        SPK_LAUNCHED();
        launch_fused_stdp(dim3(unsigned(cout), unsigned(cdiv(int64_t(cin) * kh * kw, STDP_THREADS))), s, kernel,
                          bits.p, wkey.p, stdp, *cfg);
        SPK_LAUNCHED();
    }
}

// Makes a plan; shapes: host [n][3] of (t, h, w).
LayerPlan *make_plan(int64_t n, const int32_t *shapes, int64_t cin, int64_t cout, int64_t kh, int64_t kw, int64_t pt,
                     int64_t pl, int64_t pb, int64_t pr, double threshold, int64_t ry, int64_t rx, int64_t count,
                     int64_t pkh, int64_t pkw, int64_t psh, int64_t psw, bool tune, const uint8_t *in_ptr_hint)
{
    SPK_ARG(n >= 1 && shapes, "Fused layer: no instances.");
    SPK_ARG(n <= 65535, "Fused layer: at most 65535 instances per plan.");
    SPK_ARG(cin >= 1 && cout >= 2 && kh >= 1 && kw >= 1, "Fused layer: invalid shapes.");
    SPK_ARG(pt >= 0 && pl >= 0 && pb >= 0 && pr >= 0, "Pad items are incorrect.");
    SPK_UNSUP(cout <= 256, "Fused layer: at most 256 output channels.");
    SPK_UNSUP(threshold >= 0.0, "Fused layer: the conv threshold must be >= 0.");
    SPK_UNSUP(threshold_scales(threshold), "Fused layer: the threshold is too small to scale exactly.");
    SPK_ARG(ry >= 0 && rx >= 0 && count >= 1, "Fused layer: invalid convwta parameters.");
    SPK_UNSUP(count <= SEL_MAX, "Fused layer: at most " << SEL_MAX << " winners per instance on the region path.");
    SPK_UNSUP(kh + RW + kw <= 64, "Fused layer: kernel too large for the region path.");
    auto P = std::make_unique<LayerPlan>();
    SPK_HIP(hipGetDevice(&P->device));
    P->n = n, P->cin = int32_t(cin), P->cout = int32_t(cout), P->kh = int32_t(kh), P->kw = int32_t(kw);
    P->pt = int32_t(pt), P->pl = int32_t(pl), P->count = int32_t(count), P->ry = int32_t(ry), P->rx = int32_t(rx);
    P->th = float(threshold), P->tune = tune;
    P->pool = pkh > 0;
    if (P->pool)
    {
        SPK_ARG(pkh >= 1 && pkw >= 1 && psh >= 1 && psw >= 1, "Pool: kernel and stride must be positive.");
        SPK_UNSUP(pkh * pkw <= 16, "Fused layer: pool windows of at most 16 pixels.");
        P->pkh = int32_t(pkh), P->pkw = int32_t(pkw), P->psh = int32_t(psh), P->psw = int32_t(psw);
    }
    const int64_t cin_pad = use_pair(cin, kh, kw) ? 4 : cdiv(cin, 8) * 8;
    P->host.resize(size_t(n));
    int64_t in_b = 0, out_b = 0, pool_b = 0, pix = 0, fs_b = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        const int64_t t = shapes[3 * i], h = shapes[3 * i + 1], w = shapes[3 * i + 2];
        SPK_ARG(t >= 1, "Fused layer: every instance needs at least one time step.");
        SPK_UNSUP(t <= 32, "Fused layer: at most 32 time steps per instance.");
        SPK_ARG(h >= 1 && w >= 1, "Fused layer: invalid instance shape.");
        const int64_t ho = h + pt + pb - kh + 1, wo = w + pl + pr - kw + 1;
        SPK_ARG(ho >= 1 && wo >= 1, "Input and kernel shapes are not compatible.");
        SPK_UNSUP(ho * wo <= 65536, "Fused layer: at most 65536 output pixels per instance.");
        // the buffer staging's 32-bit offsets run up to one step past the input
        SPK_UNSUP((t + 1) * cin * h * w < (int64_t(1) << 31), "Fused layer: an instance's input must be under 2 GiB.");
        int64_t pho = 0, pwo = 0;
        if (P->pool)
        {
            SPK_ARG(ho >= pkh && wo >= pkw, "Pool: input and kernel shapes are not compatible.");
            pho = (ho - pkh) / psh + 1, pwo = (wo - pkw) / psw + 1;
        }
        RegionInst &I = P->host[size_t(i)];
        I.in_off = in_b, I.out_off = out_b, I.pool_off = pool_b, I.sum_off = pix, I.fs_off = fs_b;
        I.t = int32_t(t), I.h = int32_t(h), I.w = int32_t(w), I.ho = int32_t(ho), I.wo = int32_t(wo);
        I.pho = int32_t(pho), I.pwo = int32_t(pwo), I.cols = int32_t(cdiv(wo, RW));
        in_b += t * cin * h * w, out_b += t * cout * ho * wo, pool_b += t * cout * pho * pwo;
        pix += ho * wo, fs_b += h * w * cin_pad;
        P->tmax = std::max(P->tmax, int32_t(t));
        P->max_hw = std::max(P->max_hw, ho * wo), P->max_phw = std::max(P->max_phw, pho * pwo);
    }
    P->in_bytes = in_b, P->out_bytes = out_b, P->pool_bytes = pool_b, P->pixels = pix, P->fs_bytes = fs_b;
    // dword staging: every row a whole number of aligned dwords (then every
    // instance's packed input starts 4-aligned too)
    bool fast = true;
    for (auto &I : P->host) fast = fast && I.w % 4 == 0;
    fast = fast && (reinterpret_cast<uintptr_t>(in_ptr_hint) & 3) == 0;
    P->fast = fast;

    // candidates: region rows nct * m, m = 1, 2, 4, 8 where the LDS holds them
    int cus = 0;
    SPK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device));
    for (int32_t m : {1, 2, 4, 8})
    {
        RegionGeom g;
        size_t lds = 0;
        if (!region_config(*P, m, fast, g, lds)) continue;
        std::vector<int2> items;
        for (int64_t i = 0; i < n; ++i)
        {
            const RegionInst &I = P->host[size_t(i)];
            const int64_t regions = cdiv(I.ho, g.rh) * I.cols;
            for (int64_t r = 0; r < regions; ++r) items.push_back(make_int2(int32_t(i), int32_t(r)));
        }
        SPK_ARG(items.size() < (1ull << 31), "Fused layer: too many regions.");
        g.items = int64_t(items.size());
        auto tab = std::make_unique<DeviceBuffer<int2>>(g.items);
        SPK_HIP(hipMemcpy(tab->p, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
        auto rc = std::make_unique<RegionConfig>();
        rc->g = g, rc->lds = lds, rc->items = tab->p;
        rc->grid = int(std::min<int64_t>(g.items, cus));
        // chunk cuts at instance boundaries, about g.items / PIPE_CHUNKS regions each
        rc->cut_item.push_back(0), rc->cut_inst.push_back(0);
        int64_t acc = 0;
        for (int64_t i = 0; i < n; ++i)
        {
            acc += cdiv(P->host[size_t(i)].ho, g.rh) * P->host[size_t(i)].cols;
            const int64_t c = int64_t(rc->cut_item.size());
            if (i + 1 < n && c < PIPE_CHUNKS && acc * PIPE_CHUNKS >= c * g.items)
                rc->cut_item.push_back(acc), rc->cut_inst.push_back(i + 1);
        }
        rc->cut_item.push_back(g.items), rc->cut_inst.push_back(n);
        P->item_tables.push_back(std::move(tab));
        P->configs.push_back(std::move(rc));
        if (!tune) break;  // without tuning: the smallest region that fits
    }
    SPK_UNSUP(!P->configs.empty(), "Fused layer: the region stage does not fit the 160 KiB LDS (T * Cin too large).");
    const RegionGeom &g = P->configs[0]->g;
    SPK_UNSUP(g.steps <= MAX_STEPS, "Fused layer: Cin*Kh*Kw too large (K <= 512 after padding).");
    // without tuning: outputs written inside the region kernel when a step's
    // MFMA chain (2 * steps MFMAs of 16 cycles a wave) outlasts the issue of
    // its 16-byte output segments (about one a cycle for cout * nct segments
    // a block), else by the separate writers, whose stores coalesce
    P->fuse_out = 32 * g.steps >= g.cout * g.nct;

    P->inst.reset(n);
    SPK_HIP(hipMemcpy(P->inst.p, P->host.data(), size_t(n) * sizeof(RegionInst), hipMemcpyHostToDevice));
    P->frag.reset(int64_t(g.cg) * 2 * g.steps * 64);
    P->summary.reset(pix);
    P->fs.reset(fs_b);
    P->wkey.reset(n * count);
    P->imax.reset(n);
    const int64_t ksz = cin * kh * kw;
    P->stdp = FusedStdp{n * count, int32_t(count), int32_t(cin), int32_t(cin_pad), 0, 0, int32_t(kh), int32_t(kw),
                        0, 0, 0, 0, 0, int32_t(cdiv(ksz, 32))};
    SPK_ARG(n * count < (1LL << 31), "Fused layer: too many winner slots.");
    P->bits.reset(n * count * P->stdp.words);
    return P.release();
}

namespace
{
// The plans of the uniform path, kept per (shape, stream, device) so repeated
// calls allocate nothing; a stream's calls are ordered, so its plan's
// buffers are never in use by two calls at once.
struct PlanKey
{
    std::vector<int64_t> v;
    bool operator==(const PlanKey &o) const { return v == o.v; }
};
struct CachedPlan
{
    PlanKey key;
    std::unique_ptr<LayerPlan> plan;
    bool captured = false;  // a stream capture has recorded a run of it
};
// A plan that a capture has used is baked into the captured graph (its tables
// and work buffers are kernel arguments), so it is never destroyed: eviction
// and cache_clear move it to `retired`, kept until the process ends.
struct PlanCache
{
    std::mutex mu;
    std::vector<CachedPlan> plans;  // most recent last
    std::vector<std::unique_ptr<LayerPlan>> retired;
};
PlanCache &plan_cache()
{
    static PlanCache *c = new PlanCache;
    return *c;
}
constexpr size_t PLAN_CACHE = 4;
}  // namespace

void clear_plan_cache()
{
    PlanCache &c = plan_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    if (!c.plans.empty()) SPK_HIP(hipDeviceSynchronize());
    for (auto &e : c.plans)
        if (e.captured) c.retired.push_back(std::move(e.plan));
    c.plans.clear();
}

// Uniform batches whose sample does not fit the LDS whole (spk_stdp_layer_fused).
void region_layer_step(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, float *kernel,
                       int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                       double threshold, int64_t ry, int64_t rx, int64_t count, const spk_stdp_config *cfg,
                       uint8_t *out, uint8_t *pooled, int64_t pkh, int64_t pkw, int64_t psh, int64_t psw,
                       spk_winner *winners, int64_t *counts, float *pot, hipStream_t s)
{
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    int64_t thbits = 0;
    std::memcpy(&thbits, &threshold, sizeof(thbits));
    const int64_t fast = (reinterpret_cast<uintptr_t>(in) & 3) == 0;
    PlanKey key{{n, t, cin, h, w, cout, kh, kw, pt, pl, pb, pr, thbits, ry, rx, count, pooled ? pkh : 0, pkw, psh,
                 psw, fast, int64_t(reinterpret_cast<uintptr_t>(s)), dev}};
    PlanCache &c = plan_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SPK_HIP(hipStreamIsCapturing(s, &cap));
    CachedPlan *E = nullptr;
    for (size_t i = 0; i < c.plans.size(); ++i)
        if (c.plans[i].key == key)
        {
            std::rotate(c.plans.begin() + long(i), c.plans.begin() + long(i) + 1, c.plans.end());
            E = &c.plans.back();
            break;
        }
    if (!E)
    {
        SPK_ARG(cap == hipStreamCaptureStatusNone,
                "Fused layer: this shape's plan must be made, which a stream capture cannot do: run the call once on "
                "the capturing stream before capturing it.");
        std::vector<int32_t> shapes(size_t(3 * n));
        for (int64_t i = 0; i < n; ++i)
            shapes[size_t(3 * i)] = int32_t(t), shapes[size_t(3 * i + 1)] = int32_t(h),
                               shapes[size_t(3 * i + 2)] = int32_t(w);
        std::unique_ptr<LayerPlan> made(make_plan(n, shapes.data(), cin, cout, kh, kw, pt, pl, pb, pr, threshold, ry,
                                                  rx, count, pooled ? pkh : 0, pkw, psh, psw, false, in));
        SPK_ARG(made->in_bytes == n * t * cin * h * w, "Fused layer: input is not packed contiguously.");
        if (c.plans.size() == PLAN_CACHE)
        {
            SPK_HIP(hipDeviceSynchronize());  // the oldest plan's buffers may still be in use
            if (c.plans.front().captured) c.retired.push_back(std::move(c.plans.front().plan));
            c.plans.erase(c.plans.begin());
        }
        c.plans.push_back(CachedPlan{std::move(key), std::move(made), false});
        E = &c.plans.back();
    }
    if (cap != hipStreamCaptureStatusNone) E->captured = true;
    LayerPlan *P = E->plan.get();
    P->run(in, kernel, cfg, out, pooled, winners, counts, pot, s);
}
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_layer_plan_create(spk_layer_plan **plan, int64_t n, const int32_t *shapes, int64_t cin, int64_t cout,
                              int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                              double threshold, int64_t ry, int64_t rx, int64_t count, int64_t pkh, int64_t pkw,
                              int64_t psh, int64_t psw, int32_t tune)
{
    return guard([&] {
        SPK_ARG(plan, "Plan pointer is null.");
        *plan = nullptr;
        LayerPlan *P = make_plan(n, shapes, cin, cout, kh, kw, pt, pl, pb, pr, threshold, ry, rx, count, pkh, pkw,
                                 psh, psw, tune != 0, nullptr);
        *plan = reinterpret_cast<spk_layer_plan *>(P);
    });
}

int32_t spk_layer_plan_layout(const spk_layer_plan *handle, int64_t *in_off, int64_t *out_off, int64_t *pool_off,
                              int64_t *totals)
{
    return guard([&] {
        SPK_ARG(handle, "Plan is null.");
        const LayerPlan *plan = reinterpret_cast<const LayerPlan *>(handle);
        for (int64_t i = 0; i < plan->n; ++i)
        {
            const RegionInst &I = plan->host[size_t(i)];
            if (in_off) in_off[i] = I.in_off;
            if (out_off) out_off[i] = I.out_off;
            if (pool_off) pool_off[i] = I.pool_off;
        }
        if (totals) totals[0] = plan->in_bytes, totals[1] = plan->out_bytes, totals[2] = plan->pool_bytes;
    });
}

int32_t spk_layer_plan_run(spk_layer_plan *handle, const void *in, void *kernel, int32_t kernel_type,
                           const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled,
                           spk_winner *winners, int64_t *counts, void *stream)
{
    return guard([&] {
        SPK_ARG(handle, "Plan is null.");
        LayerPlan *plan = reinterpret_cast<LayerPlan *>(handle);
        SPK_REQUIRE(kernel_type == SPK_F32, SPK_ERR_UNSUPPORTED,
                    "Fused layer: kernel data type " << type_name(kernel_type) << " is not supported.");
        SPK_ARG(in && kernel && winners && counts, "Fused layer: input, kernel and winner buffers are required.");
        SPK_ARG(!plan->fast || (reinterpret_cast<uintptr_t>(in) & 3) == 0,
                "Fused layer: the input must be 4-byte aligned.");
        SPK_ARG(!pooled || plan->pool, "Fused layer: the plan was made without pooling.");
        int dev = 0;
        SPK_HIP(hipGetDevice(&dev));
        SPK_ARG(dev == plan->device, "Fused layer: the plan belongs to another device.");
        if (plan->tune && !plan->tuned)
        {
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            SPK_HIP(hipStreamIsCapturing(S(stream), &cap));
            SPK_ARG(cap == hipStreamCaptureStatusNone,
                    "Fused layer: a tuning plan times its candidates on its first run, which a stream capture cannot "
                    "do: run it once before capturing it.");
        }
        const StdpConfigs cfg = nconfigs > 0 ? pack_configs(configs, nconfigs) : StdpConfigs{};
        plan->run((const uint8_t *)in, (float *)kernel, nconfigs > 0 ? &cfg.c[0] : nullptr, (uint8_t *)out,
                  (uint8_t *)pooled, winners, counts, nullptr, S(stream));
    });
}

int32_t spk_layer_plan_candidates(const spk_layer_plan *handle, int32_t *rows, int32_t *chunks, int32_t capacity,
                                  int32_t *count, int32_t *chosen, int32_t *writes)
{
    return guard([&] {
        SPK_ARG(handle, "Plan is null.");
        const LayerPlan *plan = reinterpret_cast<const LayerPlan *>(handle);
        const int32_t k = int32_t(plan->configs.size());
        for (int32_t i = 0; i < k && i < capacity; ++i)
        {
            const RegionGeom &g = plan->configs[size_t(i)]->g;
            if (rows) rows[i] = g.rh;
            if (chunks) chunks[i] = int32_t(cdiv(plan->tmax, g.tc));
        }
        if (count) *count = k;
        if (chosen) *chosen = plan->chosen;
        if (writes) *writes = plan->fuse_out ? 1 : 0;
    });
}

int32_t spk_layer_plan_select(spk_layer_plan *handle, int32_t index, int32_t writes)
{
    return guard([&] {
        SPK_ARG(handle, "Plan is null.");
        LayerPlan *plan = reinterpret_cast<LayerPlan *>(handle);
        SPK_ARG(index >= 0 && index < int32_t(plan->configs.size()), "Plan candidate index out of range.");
        SPK_ARG(writes >= -1 && writes <= 1, "Plan write mode must be -1, 0 or 1.");
        plan->chosen = index;
        if (writes >= 0) plan->fuse_out = writes == 1;
        plan->tuned = true;
    });
}

int32_t spk_layer_plan_destroy(spk_layer_plan *handle)
{
    return guard([&] {
        if (!handle) return;
        SPK_HIP(hipDeviceSynchronize());
        delete reinterpret_cast<LayerPlan *>(handle);
    });
}
}
