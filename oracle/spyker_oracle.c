/*
 * CPU oracle for the Spyker dense STDP layer path (C11 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product never links it.
 *
 * Each function restates the reference CPU function cited beside it.  Where
 * the reference is compiled with FMA contraction (-march=native on an FMA
 * host, GCC's default -ffp-contract=fast), the contracted form is written
 * out with an explicit fma() so the result does not depend on how this file
 * is compiled.  The file is built with -ffp-contract=off for the same reason.
 *
 * Parity status of each function (what pins it) is listed in the header of
 * tests/test_oracle.py.
 */
#include "spyker_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1e-5f /* Epsilon, src/spyker/base.h:776 */

static inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }

/* float -> U16 as GCC on x86-64 emits it (src/spyker/cpu/code.cpp:268,316
 * store F32 into U16): cvttss2si to a 32-bit integer, where NaN and values
 * out of range give 0x80000000, then keep the low 16 bits. */
static inline uint16_t f32_to_u16(float f)
{
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return 0;
    return (uint16_t)(uint32_t)(int32_t)f;
}

/* ------------------------------------------------------------------------ */
/* Rank-order coding                                                        */
/* ------------------------------------------------------------------------ */

typedef struct
{
    float v;
    uint32_t i;
} keyed;

/* Order: value descending, index ascending.  The reference uses std::sort
 * (unstable) with comp = input(i) > input(j) (src/spyker/cpu/code.cpp:314);
 * ties therefore occupy one contiguous block of ranks whose internal order is
 * unspecified there and index-ascending here. */
static int keyed_cmp(const void *a, const void *b)
{
    const keyed *x = (const keyed *)a, *y = (const keyed *)b;
    if (x->v > y->v) return -1;
    if (x->v < y->v) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* src/spyker/cpu/code.cpp:301-317 (sorted) */
void orc_rank_times_sort_f32(const float *in, int64_t len, int64_t time, uint16_t *times)
{
    keyed *k = (keyed *)malloc(sizeof(keyed) * (size_t)len);
    for (int64_t i = 0; i < len; ++i) k[i].v = in[i], k[i].i = (uint32_t)i;
    qsort(k, (size_t)len, sizeof(keyed), keyed_cmp);
    /* lower_bound of the last (smallest) element under comp: the number of
     * elements strictly greater than the minimum (code.cpp:302-307). */
    float last = k[len - 1].v;
    int64_t lo = 0, hi = len;
    while (lo < hi)
    {
        int64_t mid = lo + (hi - lo) / 2;
        if (k[mid].v > last)
            lo = mid + 1;
        else
            hi = mid;
    }
    float scale = ((float)time + EPS) / (float)lo;
    for (int64_t i = 0; i < len; ++i) times[k[i].i] = f32_to_u16((float)i * scale);
    free(k);
}

/* src/spyker/cpu/code.cpp:271-282: spike[t] = (time <= t), cumulative. */
static void times_to_spikes(const uint16_t *times, int64_t len, int64_t time, uint8_t *out)
{
    for (int64_t t = 0; t < time; ++t)
    {
        uint16_t max = (uint16_t)t;
        uint8_t *od = out + t * len;
        for (int64_t j = 0; j < len; ++j) od[j] = times[j] <= max;
    }
}

/* src/spyker/cpu/code.cpp:319-335 */
void orc_rank_code_sort_f32(const float *in, int64_t n, int64_t len, int64_t time, uint8_t *out, int32_t threads)
{
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic)
    for (int64_t i = 0; i < n; ++i)
    {
        uint16_t *times = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)len);
        orc_rank_times_sort_f32(in + i * len, len, time, times);
        times_to_spikes(times, len, time, out + i * time * len);
        free(times);
    }
}

/* src/spyker/cpu/code.cpp:43-55 (latency times of one sample); the
 * subtraction from `end` is contracted into one fused multiply-add in the
 * reference build. */
static void latency_times(const float *x, int64_t len, int64_t time, uint16_t *times)
{
    float mn = x[0], mx = x[0];
    for (int64_t i = 0; i < len; ++i)
    {
        mn = x[i] < mn ? x[i] : mn; /* std::min(a,b) = b < a ? b : a */
        mx = mx < x[i] ? x[i] : mx; /* std::max(a,b) = a < b ? b : a */
    }
    float end = (float)time + EPS, scale = end / (mx - mn);
    for (int64_t i = 0; i < len; ++i) times[i] = f32_to_u16(fmaf(-scale, x[i] - mn, end));
}

/* src/spyker/cpu/code.cpp:258-269 (latency) */
void orc_rank_code_latency_f32(const float *in, int64_t n, int64_t len, int64_t time, uint8_t *out)
{
    uint16_t *times = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)len);
    for (int64_t s = 0; s < n; ++s)
    {
        latency_times(in + s * len, len, time, times);
        times_to_spikes(times, len, time, out + s * time * len);
    }
    free(times);
}

/* ------------------------------------------------------------------------ */
/* Rate (Poisson) coding                                                    */
/* ------------------------------------------------------------------------ */

/* MT19937 (Matsumoto & Nishimura 1998), the std::mt19937 the reference draws
 * from (code.cpp:140 `using Random = std::mt19937`): 32-bit words, n = 624,
 * m = 397, seeded as the standard's seed(value). */
typedef struct
{
    uint32_t x[624];
    int p;
} orc_mt;

void orc_mt_seed(orc_mt *g, uint32_t seed)
{
    g->x[0] = seed;
    for (int i = 1; i < 624; ++i) g->x[i] = 1812433253u * (g->x[i - 1] ^ (g->x[i - 1] >> 30)) + (uint32_t)i;
    g->p = 624;
}

uint32_t orc_mt_next(orc_mt *g)
{
    if (g->p >= 624)
    {
        for (int k = 0; k < 624; ++k)
        {
            uint32_t y = (g->x[k] & 0x80000000u) | (g->x[(k + 1) % 624] & 0x7fffffffu);
            g->x[k] = g->x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->p = 0;
    }
    uint32_t z = g->x[g->p++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
}

/* The first n outputs of std::mt19937(seed). */
void orc_mt_stream(uint32_t seed, int64_t n, uint32_t *out)
{
    orc_mt g;
    orc_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_mt_next(&g);
}

/* code.cpp:168-191 for one sample: with rows[j] the spike-time row of
 * element j, draw(j) = table[rows[j]][next() & (bins - 1)]; every element
 * draws once in order, then at each step t, in element order, an element
 * whose next time equals t counts a spike and draws its next interval; the
 * output at step t is the count so far.  times/values are U16 as there. */
static void poisson_sample(const uint16_t *rows, int64_t len, int64_t time, const uint16_t *table, int64_t bins,
                           orc_mt *g, uint16_t *times, uint16_t *values, uint8_t *out)
{
    for (int64_t j = 0; j < len; ++j) values[j] = 0;
    for (int64_t j = 0; j < len; ++j) times[j] = table[rows[j] * bins + (orc_mt_next(g) & (uint32_t)(bins - 1))];
    for (int64_t t = 0; t < time; ++t)
    {
        for (int64_t j = 0; j < len; ++j)
            if (times[j] == t)
            {
                values[j] += 1;
                times[j] = (uint16_t)(times[j] + table[rows[j] * bins + (orc_mt_next(g) & (uint32_t)(bins - 1))]);
            }
        for (int64_t j = 0; j < len; ++j) out[t * len + j] = (uint8_t)values[j];
    }
}

/* code.cpp:192-230 (poisson, poissort) with the table given: sample i draws
 * from std::mt19937(rnd + i) (rnd = the shared generator's draw for the
 * call), its rows are the rank-code times (latency or sorted).  Every row
 * index must be below `rows`; -1 is returned otherwise. */
int32_t orc_rate_code_f32(const float *in, int64_t n, int64_t len, int64_t time, int32_t sort, uint64_t rnd,
                          const uint16_t *table, int64_t rows, int64_t bins, uint8_t *out)
{
    uint16_t *r = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)len * 3);
    uint16_t *times = r + len, *values = r + 2 * len;
    int32_t rc = 0;
    for (int64_t i = 0; i < n && rc == 0; ++i)
    {
        if (sort)
            orc_rank_times_sort_f32(in + i * len, len, time, r);
        else
            latency_times(in + i * len, len, time, r);
        for (int64_t j = 0; j < len; ++j)
            if (r[j] >= rows) rc = -1;
        if (rc) break;
        orc_mt g;
        orc_mt_seed(&g, (uint32_t)(rnd + (uint64_t)i));
        poisson_sample(r, len, time, table, bins, &g, times, values, out + i * time * len);
    }
    free(r);
    return rc;
}

/* src/spyker/cpu/fire.cpp:84-114 (rate fire): per element, over the steps,
 * a spike is counted when value + threshold < x, and value becomes x; the
 * output at step t is the count so far. */
void orc_rate_fire_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, uint8_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < len; ++j)
        {
            float value = 0.0f;
            uint8_t spikes = 0;
            for (int64_t s = 0; s < t; ++s)
            {
                const float x = in[(i * t + s) * len + j];
                if (value + threshold < x) ++spikes, value = x;
                out[(i * t + s) * len + j] = spikes;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* threshold / fire                                                         */
/* ------------------------------------------------------------------------ */

/* src/spyker/cpu/fire.cpp:41-46 */
void orc_threshold_f32(float *x, int64_t n, float threshold, float value)
{
    for (int64_t i = 0; i < n; ++i)
        if (x[i] <= threshold) x[i] = value;
}

/* src/spyker/cpu/fire.cpp:67-71 */
void orc_rank_fire_f32(const float *x, uint8_t *out, int64_t n, float threshold)
{
    for (int64_t i = 0; i < n; ++i) out[i] = x[i] > threshold;
}

/* ------------------------------------------------------------------------ */
/* inhibit                                                                  */
/* ------------------------------------------------------------------------ */

/* src/spyker/cpu/inhibit.cpp:41-121.  Per sample and pixel: m_t, ch_t = the
 * first channel maximum at step t (strict >, starting from channel 0); cnt =
 * #{t : m_t > threshold}; the kept channel is ch at step T - max(cnt, 1);
 * every value is multiplied by (channel == kept). */
void orc_rank_inhibit_f32(float *x, int64_t n, int64_t t, int64_t c, int64_t hw, float threshold, int32_t threads)
{
#pragma omp parallel for num_threads(threads > 0 ? threads : 1)
    for (int64_t s = 0; s < n; ++s)
    {
        float *xs = x + s * t * c * hw;
        uint16_t *ch = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(t * hw));
        uint16_t *sum = (uint16_t *)calloc((size_t)hw, sizeof(uint16_t));
        for (int64_t i = 0; i < t; ++i)
            for (int64_t p = 0; p < hw; ++p)
            {
                float m = xs[(i * c) * hw + p];
                uint16_t a = 0;
                for (int64_t j = 1; j < c; ++j)
                {
                    float v = xs[(i * c + j) * hw + p];
                    if (v > m) m = v, a = (uint16_t)j;
                }
                ch[i * hw + p] = a;
                sum[p] = (uint16_t)(sum[p] + (m > threshold));
            }
        for (int64_t p = 0; p < hw; ++p)
        {
            uint16_t cnt = sum[p] > 1 ? sum[p] : 1;
            uint16_t keep = ch[(int64_t)(uint16_t)(t - cnt) * hw + p];
            for (int64_t i = 0; i < t; ++i)
                for (int64_t j = 0; j < c; ++j) xs[(i * c + j) * hw + p] *= (float)((uint16_t)j == keep);
        }
        free(ch);
        free(sum);
    }
}

/* ------------------------------------------------------------------------ */
/* winner-take-all                                                          */
/* ------------------------------------------------------------------------ */

/* src/spyker/cpu/wta.cpp:49-64: value of the earliest step above threshold. */
static void wta_values_f32(const float *x, int64_t t, int64_t len, float threshold, float *temp)
{
    float mn = x[0], mx = x[0];
    for (int64_t i = 0; i < t * len; ++i)
    {
        mn = x[i] < mn ? x[i] : mn;
        mx = mx < x[i] ? x[i] : mx;
    }
    float diff = (mx - mn) * 2.0f;
    for (int64_t j = 0; j < len; ++j) temp[j] = -FLT_MAX;
    for (int64_t i = t - 1; i >= 0; --i)
    {
        const float *xd = x + i * len;
        float level = (float)(t - i - 1);
        for (int64_t j = 0; j < len; ++j)
            if (xd[j] > threshold) temp[j] = level + (xd[j] - mn) / diff;
    }
}

/* std::max_element: first maximum under operator<. */
static int64_t first_max_f32(const float *v, int64_t len)
{
    int64_t best = 0;
    for (int64_t i = 1; i < len; ++i)
        if (v[best] < v[i]) best = i;
    return best;
}

/* src/spyker/cpu/wta.cpp:128-150 */
static int conv_select_f32(float *temp, int64_t c, int64_t h, int64_t w, int64_t ry, int64_t rx, orc_winner *win)
{
    int64_t len = h * w;
    int64_t idx = first_max_f32(temp, c * len);
    if (fabsf(temp[idx] - (-FLT_MAX)) < EPS) return 0;
    win->c = 0, win->t = 0;
    win->z = idx / len, win->y = (idx % len) / w, win->x = (idx % len) % w;
    int64_t sj = imax(0, win->y - ry), ej = imin(h, win->y + ry + 1);
    int64_t sk = imax(0, win->x - rx), ek = imin(w, win->x + rx + 1);
    for (int64_t i = 0; i < c; ++i)
        for (int64_t j = sj; j < ej; ++j)
            for (int64_t k = sk; k < ek; ++k) temp[(i * h + j) * w + k] = -FLT_MAX;
    for (int64_t j = 0; j < len; ++j) temp[win->z * len + j] = -FLT_MAX;
    return 1;
}

/* src/spyker/cpu/wta.cpp:152-182 */
void orc_rank_convwta_f32(const float *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                          int64_t rx, int64_t count, float threshold, orc_winner *winners, int64_t *counts,
                          int32_t threads)
{
    int64_t len = c * h * w;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1)
    for (int64_t s = 0; s < n; ++s)
    {
        float *temp = (float *)malloc(sizeof(float) * (size_t)len);
        wta_values_f32(x + s * t * len, t, len, threshold, temp);
        int64_t k = 0;
        for (; k < count; ++k)
            if (!conv_select_f32(temp, c, h, w, ry, rx, &winners[s * count + k])) break;
        counts[s] = k;
        free(temp);
    }
}

/* Integer branch, src/spyker/cpu/wta.cpp:66-85 (threshold unused there). */
void orc_rank_convwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                         int64_t rx, int64_t count, orc_winner *winners, int64_t *counts)
{
    int64_t len = c * h * w, hw = h * w;
    uint16_t *temp = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)len);
    for (int64_t s = 0; s < n; ++s)
    {
        const uint8_t *xs = x + s * t * len;
        memset(temp, 0, sizeof(uint16_t) * (size_t)len);
        for (int64_t i = t - 1; i >= 0; --i)
            for (int64_t j = 0; j < len; ++j)
                if (xs[i * len + j]) temp[j] = (uint16_t)(t - i);
        int64_t k = 0;
        for (; k < count; ++k)
        {
            int64_t idx = 0;
            for (int64_t j = 1; j < len; ++j)
                if (temp[idx] < temp[j]) idx = j;
            if (temp[idx] == 0) break;
            orc_winner *win = &winners[s * count + k];
            win->c = 0, win->t = 0;
            win->z = idx / hw, win->y = (idx % hw) / w, win->x = (idx % hw) % w;
            int64_t sj = imax(0, win->y - ry), ej = imin(h, win->y + ry + 1);
            int64_t sk = imax(0, win->x - rx), ek = imin(w, win->x + rx + 1);
            for (int64_t i = 0; i < c; ++i)
                for (int64_t j = sj; j < ej; ++j)
                    for (int64_t q = sk; q < ek; ++q) temp[(i * h + j) * w + q] = 0;
            for (int64_t j = 0; j < hw; ++j) temp[win->z * hw + j] = 0;
        }
        counts[s] = k;
    }
    free(temp);
}

/* src/spyker/cpu/wta.cpp:87-126 */
void orc_rank_fcwta_f32(const float *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                        float threshold, orc_winner *winners, int64_t *counts)
{
    float *temp = (float *)malloc(sizeof(float) * (size_t)f);
    for (int64_t s = 0; s < n; ++s)
    {
        wta_values_f32(x + s * t * f, t, f, threshold, temp);
        int64_t k = 0;
        for (; k < count; ++k)
        {
            int64_t idx = first_max_f32(temp, f);
            if (fabsf(temp[idx] - (-FLT_MAX)) < EPS) break;
            orc_winner *win = &winners[s * count + k];
            win->c = win->t = win->z = win->y = 0;
            win->x = idx;
            int64_t a = imax(0, idx - radius), b = imin(f, idx + radius + 1);
            for (int64_t j = a; j < b; ++j) temp[j] = -FLT_MAX;
        }
        counts[s] = k;
    }
    free(temp);
}

/* ------------------------------------------------------------------------ */
/* pad / pool                                                               */
/* ------------------------------------------------------------------------ */

/* src/spyker/cpu/pool.cpp:41-51 */
void orc_pad_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t pt, int64_t pl,
                int64_t pb, int64_t pr, uint8_t value)
{
    int64_t oh = h + pt + pb, ow = w + pl + pr;
    for (int64_t i = 0; i < planes; ++i)
    {
        memset(out + i * oh * ow, value, (size_t)(oh * ow));
        for (int64_t j = 0; j < h; ++j) memcpy(out + (i * oh + j + pt) * ow + pl, in + (i * h + j) * w, (size_t)w);
    }
}

void orc_pad_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t pt, int64_t pl, int64_t pb,
                 int64_t pr, float value)
{
    int64_t oh = h + pt + pb, ow = w + pl + pr;
    for (int64_t i = 0; i < planes; ++i)
    {
        for (int64_t j = 0; j < oh * ow; ++j) out[i * oh * ow + j] = value;
        for (int64_t j = 0; j < h; ++j)
            memcpy(out + (i * oh + j + pt) * ow + pl, in + (i * h + j) * w, (size_t)w * sizeof(float));
    }
}

/* src/spyker/cpu/pool.cpp:53-86: zero pad, window max seeded with the type's
 * lowest value. */
#define POOL_BODY(T, LOWEST)                                                                           \
    int64_t ph = h + pt + pb, pw = w + pl + pr;                                                        \
    int64_t oh = (ph - kh) / sh + 1, ow = (pw - kw) / sw + 1;                                          \
    for (int64_t i = 0; i < planes; ++i)                                                               \
        for (int64_t y = 0; y < oh; ++y)                                                               \
            for (int64_t x = 0; x < ow; ++x)                                                           \
            {                                                                                          \
                T m = LOWEST;                                                                          \
                for (int64_t a = 0; a < kh; ++a)                                                       \
                    for (int64_t b = 0; b < kw; ++b)                                                   \
                    {                                                                                  \
                        int64_t iy = y * sh + a - pt, ix = x * sw + b - pl;                            \
                        T v = (iy >= 0 && iy < h && ix >= 0 && ix < w) ? in[(i * h + iy) * w + ix] : (T)0; \
                        m = m < v ? v : m;                                                             \
                    }                                                                                  \
                out[(i * oh + y) * ow + x] = m;                                                        \
            }

void orc_rank_pool_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t kh, int64_t kw,
                      int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    POOL_BODY(uint8_t, 0)
}

void orc_rank_pool_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t kh, int64_t kw,
                       int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    POOL_BODY(float, -FLT_MAX)
}

/* ------------------------------------------------------------------------ */
/* STDP                                                                     */
/* ------------------------------------------------------------------------ */

/* src/spyker/cpu/stdp.cpp:121-137: computed in F64 and rounded to the kernel
 * type; `old + rate * value` is contracted into one fused multiply-add in the
 * reference build. */
float orc_stdp_update(float w, int32_t potentiate, const orc_stdp_config *cfg)
{
    double old = (double)w;
    double value = cfg->stabilize ? (old - cfg->lower) * (cfg->upper - old) : 1.0;
    value = fma(potentiate ? cfg->positive : cfg->negative, value, old);
    double hi = value < cfg->upper ? value : cfg->upper;         /* std::min(value, upper) */
    double r = cfg->lower < hi ? hi : cfg->lower;                 /* std::max(lower, .) */
    return (float)r;
}

/* src/spyker/cpu/stdp.cpp:139-169 */
int32_t orc_rank_convstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w,
                              float *kernel, int64_t cout, int64_t kh, int64_t kw, const uint8_t *out, int64_t ho,
                              int64_t wo, const orc_stdp_config *configs, int64_t nconfigs,
                              const orc_winner *winners, const int64_t *counts, int64_t maxcount, int64_t sy,
                              int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    (void)pb, (void)pr;
    int64_t ksz = cin * kh * kw;
    uint16_t *isum = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)ksz);
    for (int64_t s = 0; s < n; ++s)
        for (int64_t k = 0; k < counts[s]; ++k)
        {
            orc_winner win = winners[s * maxcount + k];
            if (win.c < 0) continue;
            if (win.c >= nconfigs)
            {
                free(isum);
                return -1;
            }
            /* first step with a nonzero input at each tap of the padded input */
            for (int64_t c = 0; c < cin; ++c)
                for (int64_t a = 0; a < kh; ++a)
                    for (int64_t b = 0; b < kw; ++b)
                    {
                        int64_t iy = win.y * sy + a - pt, ix = win.x * sx + b - pl;
                        uint16_t first = (uint16_t)t;
                        if (iy >= 0 && iy < h && ix >= 0 && ix < w)
                            for (int64_t i = t - 1; i >= 0; --i)
                                if (in[(((s * t + i) * cin + c) * h + iy) * w + ix]) first = (uint16_t)i;
                        isum[(c * kh + a) * kw + b] = first;
                    }
            uint16_t osum = (uint16_t)t;
            for (int64_t i = t - 1; i >= 0; --i)
                if (out[(((s * t + i) * cout + win.z) * ho + win.y) * wo + win.x]) osum = (uint16_t)i;
            float *kd = kernel + win.z * ksz;
            for (int64_t i = 0; i < ksz; ++i) kd[i] = orc_stdp_update(kd[i], isum[i] <= osum, &configs[win.c]);
        }
    free(isum);
    return 0;
}

/* src/spyker/cpu/stdp.cpp:42-97 */
int32_t orc_rank_fcstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t fin, float *kernel, int64_t fout,
                            const uint8_t *out, const orc_stdp_config *configs, int64_t nconfigs,
                            const orc_winner *winners, const int64_t *counts, int64_t maxcount)
{
    uint16_t *isum = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)fin);
    for (int64_t s = 0; s < n; ++s)
    {
        for (int64_t f = 0; f < fin; ++f) isum[f] = (uint16_t)t;
        for (int64_t i = t - 1; i >= 0; --i)
            for (int64_t f = 0; f < fin; ++f)
                if (in[(s * t + i) * fin + f]) isum[f] = (uint16_t)i;
        for (int64_t k = 0; k < counts[s]; ++k)
        {
            orc_winner win = winners[s * maxcount + k];
            if (win.c < 0) continue;
            if (win.c >= nconfigs)
            {
                free(isum);
                return -1;
            }
            uint16_t osum = (uint16_t)t;
            for (int64_t i = t - 1; i >= 0; --i)
                if (out[(s * t + i) * fout + win.x]) osum = (uint16_t)i;
            float *kd = kernel + win.x * fin;
            for (int64_t f = 0; f < fin; ++f) kd[f] = orc_stdp_update(kd[f], isum[f] <= osum, &configs[win.c]);
        }
    }
    free(isum);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* contractions                                                             */
/* ------------------------------------------------------------------------ */

static inline float load_in(const void *in, int32_t is_u8, int64_t i)
{
    return is_u8 ? (float)((const uint8_t *)in)[i] : ((const float *)in)[i];
}

/* Cross-correlation with zero pad (src/spyker/cpu/conv.cpp semantics; the
 * reference's arithmetic lives in oneDNN/cuDNN, so the summation order here
 * is this project's canonical order). */
void orc_conv_f32(const void *in, int32_t is_u8, int64_t n, int64_t cin, int64_t h, int64_t w, const float *kernel,
                  int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                  int64_t pr, float *out, int32_t threads)
{
    int64_t ho = (h + pt + pb - kh) / sy + 1, wo = (w + pl + pr - kw) / sx + 1;
#pragma omp parallel for collapse(2) num_threads(threads > 0 ? threads : 1)
    for (int64_t s = 0; s < n; ++s)
        for (int64_t o = 0; o < cout; ++o)
            for (int64_t y = 0; y < ho; ++y)
                for (int64_t x = 0; x < wo; ++x)
                {
                    float acc = 0.0f;
                    for (int64_t c = 0; c < cin; ++c)
                        for (int64_t a = 0; a < kh; ++a)
                        {
                            int64_t iy = y * sy + a - pt;
                            if (iy < 0 || iy >= h) continue;
                            for (int64_t b = 0; b < kw; ++b)
                            {
                                int64_t ix = x * sx + b - pl;
                                if (ix < 0 || ix >= w) continue;
                                acc = fmaf(load_in(in, is_u8, ((s * cin + c) * h + iy) * w + ix),
                                           kernel[((o * cin + c) * kh + a) * kw + b], acc);
                            }
                        }
                    out[((s * cout + o) * ho + y) * wo + x] = acc;
                }
}

void orc_conv_f64(const void *in, int32_t is_u8, int64_t n, int64_t cin, int64_t h, int64_t w, const float *kernel,
                  int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                  int64_t pr, double *out, int32_t threads)
{
    int64_t ho = (h + pt + pb - kh) / sy + 1, wo = (w + pl + pr - kw) / sx + 1;
#pragma omp parallel for collapse(2) num_threads(threads > 0 ? threads : 1)
    for (int64_t s = 0; s < n; ++s)
        for (int64_t o = 0; o < cout; ++o)
            for (int64_t y = 0; y < ho; ++y)
                for (int64_t x = 0; x < wo; ++x)
                {
                    double acc = 0.0;
                    for (int64_t c = 0; c < cin; ++c)
                        for (int64_t a = 0; a < kh; ++a)
                        {
                            int64_t iy = y * sy + a - pt;
                            if (iy < 0 || iy >= h) continue;
                            for (int64_t b = 0; b < kw; ++b)
                            {
                                int64_t ix = x * sx + b - pl;
                                if (ix < 0 || ix >= w) continue;
                                acc += (double)load_in(in, is_u8, ((s * cin + c) * h + iy) * w + ix) *
                                       (double)kernel[((o * cin + c) * kh + a) * kw + b];
                            }
                        }
                    out[((s * cout + o) * ho + y) * wo + x] = acc;
                }
}

/* F64 kernels (opers.cpp:408-421: the input converted to F64): one fused
 * multiply-add per tap in (cin, ky, kx) order from +0, the order the device
 * conv fixes.  in_kind: 0 = U8, 1 = F32, 2 = F64. */
void orc_conv_f64k(const void *in, int32_t in_kind, int64_t n, int64_t cin, int64_t h, int64_t w,
                   const double *kernel, int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt,
                   int64_t pl, int64_t pb, int64_t pr, double *out, int32_t threads)
{
    int64_t ho = (h + pt + pb - kh) / sy + 1, wo = (w + pl + pr - kw) / sx + 1;
#pragma omp parallel for collapse(2) num_threads(threads > 0 ? threads : 1)
    for (int64_t s = 0; s < n; ++s)
        for (int64_t o = 0; o < cout; ++o)
            for (int64_t y = 0; y < ho; ++y)
                for (int64_t x = 0; x < wo; ++x)
                {
                    double acc = 0.0;
                    for (int64_t c = 0; c < cin; ++c)
                        for (int64_t a = 0; a < kh; ++a)
                        {
                            int64_t iy = y * sy + a - pt;
                            if (iy < 0 || iy >= h) continue;
                            for (int64_t b = 0; b < kw; ++b)
                            {
                                int64_t ix = x * sx + b - pl;
                                if (ix < 0 || ix >= w) continue;
                                int64_t i = ((s * cin + c) * h + iy) * w + ix;
                                double v = in_kind == 0   ? (double)((const uint8_t *)in)[i]
                                           : in_kind == 1 ? (double)((const float *)in)[i]
                                                          : ((const double *)in)[i];
                                acc = fma(v, kernel[((o * cin + c) * kh + a) * kw + b], acc);
                            }
                        }
                    out[((s * cout + o) * ho + y) * wo + x] = acc;
                }
}

/* out[r][o] = sum_f in[r][f] * k[o][f] (src/spyker/cpu/fc.cpp semantics). */
void orc_fc_f32(const void *in, int32_t is_u8, int64_t rows, int64_t fin, const float *kernel, int64_t fout,
                float *out)
{
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t o = 0; o < fout; ++o)
        {
            float acc = 0.0f;
            for (int64_t f = 0; f < fin; ++f) acc = fmaf(load_in(in, is_u8, r * fin + f), kernel[o * fin + f], acc);
            out[r * fout + o] = acc;
        }
}

/* src/spyker/cpu/dog.cpp:39-60 */
void orc_dog_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count, int64_t k,
                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads)
{
    int64_t ho = h + pt + pb - k + 1, wo = w + pl + pr - k + 1, len = count * ho * wo;
    float *mid = (float *)malloc(sizeof(float) * (size_t)(planes * 2 * len));
    orc_conv_f32(in, 0, planes, 1, h, w, kernel, 2 * count, k, k, 1, 1, pt, pl, pb, pr, mid, threads);
    for (int64_t i = 0; i < planes; ++i)
        for (int64_t j = 0; j < len; ++j)
        {
            float d = mid[(i * 2) * len + j] - mid[(i * 2 + 1) * len + j];
            out[i * len + j] = d < 0.0f ? 0.0f : d; /* std::max<T>(d, 0) */
        }
    free(mid);
}

/* src/spyker/cpu/dog.cpp:94 */
void orc_gabor_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count,
                   int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads)
{
    orc_conv_f32(in, 0, planes, 1, h, w, kernel, count, k, k, 1, 1, pt, pl, pb, pr, out, threads);
}

/* ------------------------------------------------------------------------ */
/* host kernel generators                                                   */
/* ------------------------------------------------------------------------ */

static void normalize(double *v, int64_t n)
{
    double sum = 0;
    for (int64_t i = 0; i < n; ++i) sum += v[i];
    for (int64_t i = 0; i < n; ++i) v[i] /= sum;
}

/* src/spyker/cpu/init.cpp:51-62 */
/* cpu/dog.cpp:62-84: the conv writes [planes][2][count][ho][wo]; then, per
 * plane, d = [0] - [1] gives [0] = max(d, 0) and [1] = max(-d, 0). */
void orc_log_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count, int64_t k,
                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads)
{
    int64_t ho = h + pt + pb - k + 1, wo = w + pl + pr - k + 1, len = count * ho * wo;
    orc_conv_f32(in, 0, planes, 1, h, w, kernel, 2 * count, k, k, 1, 1, pt, pl, pb, pr, out, threads);
    for (int64_t i = 0; i < planes; ++i)
        for (int64_t j = 0; j < len; ++j)
        {
            float *a = out + (i * 2) * len + j, *b = out + (i * 2 + 1) * len + j;
            float d = *a - *b, e = -d;
            *a = d < 0.0f ? 0.0f : d; /* std::max<T>(diff, 0) */
            *b = e < 0.0f ? 0.0f : e; /* std::max<T>(-diff, 0) */
        }
}

void orc_gaussian_kernel(double *kernel, int64_t k, double std)
{
    int64_t size = k / 2;
    for (int64_t i = -size; i <= size; ++i)
        for (int64_t j = -size; j <= size; ++j)
        {
            double dist = sqrt((double)(i * i + j * j));
            double r = dist / std;
            kernel[(i + size) * k + (j + size)] = exp(-.5 * (r * r));
        }
    normalize(kernel, k * k);
}

/* src/spyker/cpu/init.cpp:64-77; the sums of products are contracted into
 * fused multiply-adds in the reference build. */
void orc_gabor_kernel(double *kernel, int64_t k, double sigma, double theta, double gamma, double lambda, double psi)
{
    const double pi = 3.14159265358979323846;
    int64_t size = k / 2;
    double ct = cos(theta), st = sin(theta);
    for (int64_t i = -size; i <= size; ++i)
        for (int64_t j = -size; j <= size; ++j)
        {
            double x = fma((double)j, st, (double)i * ct);
            double y = fma((double)j, ct, (double)-i * st);
            double A = exp(-fma(x, x, gamma * gamma * y * y) / (2 * sigma * sigma));
            double B = cos(2 * pi * x / lambda + psi);
            kernel[(i + size) * k + (j + size)] = A * B;
        }
    normalize(kernel, k * k);
}

/* ------------------------------------------------------------------------ */
/* sparse event path (src/spyker/sparse/)                                    */
/* ------------------------------------------------------------------------ */
/*
 * A sparse tensor of shape [n][t][c][h][w] holds, for every (sample i, step
 * j), the list of spikes that are new at step j (base.h:493-528).  Here a
 * list set is off[n*t + 1] (list i*t+j is events [off[i*t+j], off[i*t+j+1]))
 * and events packed as z << 16 | y << 8 | x (the reference's U8 Spridx).
 * Functions that make a list set write at most `cap` events and return the
 * total, or -1 if it would not fit.
 */

#define EV(z, y, x) ((int32_t)(((uint32_t)(z) << 16) | ((uint32_t)(y) << 8) | (uint32_t)(x)))
#define EZ(e) ((int64_t)(((uint32_t)(e) >> 16) & 0xFFu))
#define EY(e) ((int64_t)(((uint32_t)(e) >> 8) & 0xFFu))
#define EX(e) ((int64_t)((uint32_t)(e)&0xFFu))

/* Lists from a first-spike-time map [n][c][h][w] (times >= t: no spike):
 * each spike goes to list (i, time) in (z, y, x) scan order
 * (sparse/code.cpp:56-66, sparse/opers.cpp:57-67). */
static int64_t lists_from_times(const uint16_t *times, int64_t n, int64_t t, int64_t len, int64_t h, int64_t w,
                                int64_t *off, int32_t *ev, int64_t cap)
{
    int64_t total = 0;
    off[0] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < t; ++j)
        {
            for (int64_t k = 0; k < len; ++k)
                if (times[i * len + k] == j)
                {
                    if (total >= cap) return -1;
                    const int64_t z = k / (h * w), y = (k / w) % h, x = k % w;
                    ev[total++] = EV(z, y, x);
                }
            off[i * t + j + 1] = total;
        }
    return total;
}

/* sparse/code.cpp:98-132: the dense rank code's spike times, as events. */
int64_t orc_sparse_code_f32(const float *in, int64_t n, int64_t c, int64_t h, int64_t w, int64_t t, int32_t sort,
                            int64_t *off, int32_t *ev, int64_t cap)
{
    const int64_t len = c * h * w;
    uint8_t *sp = (uint8_t *)malloc((size_t)(n * t * len));
    uint16_t *times = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n * len));
    if (sort)
        orc_rank_code_sort_f32(in, n, len, t, sp, 1);
    else
        orc_rank_code_latency_f32(in, n, len, t, sp);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < len; ++k)
        {
            uint16_t f = (uint16_t)t;
            for (int64_t j = t - 1; j >= 0; --j)
                if (sp[(i * t + j) * len + k]) f = (uint16_t)j;
            times[i * len + k] = f;
        }
    const int64_t r = lists_from_times(times, n, t, len, h, w, off, ev, cap);
    free(sp);
    free(times);
    return r;
}

/* SparseTensor(tensor, threshold), sparse/opers.cpp:38-80: an element's event
 * is at the first step whose value is > threshold. */
int64_t orc_sparse_from_dense_u8(const uint8_t *in, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                                 uint8_t threshold, int64_t *off, int32_t *ev, int64_t cap)
{
    const int64_t len = c * h * w;
    uint16_t *times = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n * len));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < len; ++k)
        {
            uint16_t f = (uint16_t)t;
            for (int64_t j = t - 1; j >= 0; --j)
                if (in[(i * t + j) * len + k] > threshold) f = (uint16_t)j;
            times[i * len + k] = f;
        }
    const int64_t r = lists_from_times(times, n, t, len, h, w, off, ev, cap);
    free(times);
    return r;
}

/* SparseTensor::dense(), sparse/opers.cpp:82-100: cumulative U8 spikes. */
void orc_sparse_to_dense_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                            int64_t w, uint8_t *out)
{
    const int64_t len = c * h * w;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < t; ++j)
        {
            uint8_t *o = out + (i * t + j) * len;
            if (j == 0)
                memset(o, 0, (size_t)len);
            else
                memcpy(o, o - len, (size_t)len);
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
                o[(EZ(ev[e]) * h + EY(ev[e])) * w + EX(ev[e])] = 1;
        }
}

/* Sparse::gather, sparse/opers.cpp:102-113: first spike step per element
 * (t when none), [n][c][h][w]. */
void orc_sparse_gather_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                          int64_t w, uint8_t *out)
{
    const int64_t len = c * h * w;
    for (int64_t i = 0; i < n; ++i)
    {
        uint8_t *o = out + i * len;
        memset(o, (int)t, (size_t)len);
        for (int64_t j = t - 1; j >= 0; --j)
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
                o[(EZ(ev[e]) * h + EY(ev[e])) * w + EX(ev[e])] = (uint8_t)j;
    }
}

/* sparse/pad.cpp:36-56: every event moves by (top, left). */
void orc_sparse_pad(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t pt, int64_t pl,
                    int32_t *out)
{
    for (int64_t e = 0; e < off[n * t]; ++e) out[e] = EV(EZ(ev[e]), EY(ev[e]) + pt, EX(ev[e]) + pl);
}

/* sparse/pool.cpp:36-65 (stride = kernel): an event maps to pooled cell
 * ((pt + y) / kh, (pl + x) / kw); the first event (step, then list order)
 * reaching a cell is kept, in the input's list order; cells past the
 * output are dropped. */
int64_t orc_sparse_pool(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t ho,
                        int64_t wo, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t *oof, int32_t *oev,
                        int64_t cap)
{
    uint8_t *seen = (uint8_t *)malloc((size_t)(c * ho * wo));
    int64_t total = 0;
    oof[0] = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        memset(seen, 0, (size_t)(c * ho * wo));
        for (int64_t j = 0; j < t; ++j)
        {
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
            {
                const int64_t z = EZ(ev[e]), y = (uint16_t)(pt + EY(ev[e])) / kh, x = (uint16_t)(pl + EX(ev[e])) / kw;
                if (y >= ho || x >= wo) continue;
                if (seen[(z * ho + y) * wo + x]) continue;
                seen[(z * ho + y) * wo + x] = 1;
                if (total >= cap)
                {
                    free(seen);
                    return -1;
                }
                oev[total++] = EV(z, y, x);
            }
            oof[i * t + j + 1] = total;
        }
    }
    free(seen);
    return total;
}

/* sparse/inhibit.cpp:36-66: a pixel keeps the channel of its first event
 * (step, then list order); every event of that channel there is kept. */
int64_t orc_sparse_inhibit(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                           int64_t *oof, int32_t *oev, int64_t cap)
{
    int32_t *ch = (int32_t *)malloc(sizeof(int32_t) * (size_t)(h * w));
    int64_t total = 0;
    oof[0] = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        for (int64_t p = 0; p < h * w; ++p) ch[p] = -1;
        for (int64_t e = off[i * t]; e < off[i * t + t]; ++e)
            if (ch[EY(ev[e]) * w + EX(ev[e])] < 0) ch[EY(ev[e]) * w + EX(ev[e])] = (int32_t)EZ(ev[e]);
        for (int64_t j = 0; j < t; ++j)
        {
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
                if (ch[EY(ev[e]) * w + EX(ev[e])] == EZ(ev[e]))
                {
                    if (total >= cap)
                    {
                        free(ch);
                        return -1;
                    }
                    oev[total++] = ev[e];
                }
            oof[i * t + j + 1] = total;
        }
    }
    free(ch);
    return total;
}

/* sparse/wta.cpp:40-96: events in step then list order; an event is a
 * winner unless its channel already won or it lies in the box (clipped to
 * the tensor) of an earlier winner; at most `count`. */
void orc_sparse_convwta(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                        int64_t ry, int64_t rx, int64_t count, orc_winner *winners, int64_t *counts)
{
    for (int64_t i = 0; i < n; ++i)
    {
        int64_t found = 0;
        for (int64_t j = 0; j < t && found < count; ++j)
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1] && found < count; ++e)
            {
                const int64_t z = EZ(ev[e]), y = EY(ev[e]), x = EX(ev[e]);
                int ok = 1;
                for (int64_t k = 0; k < found && ok; ++k)
                {
                    const orc_winner *b = &winners[i * count + k];
                    if (z == b->z) ok = 0;
                    const int64_t ys = b->y - ry > 0 ? b->y - ry : 0, ye = b->y + ry < h - 1 ? b->y + ry : h - 1;
                    const int64_t xs = b->x - rx > 0 ? b->x - rx : 0, xe = b->x + rx < w - 1 ? b->x + rx : w - 1;
                    if (ys <= y && y <= ye && xs <= x && x <= xe) ok = 0;
                }
                if (ok) winners[i * count + found++] = (orc_winner){0, j, z, y, x};
            }
        counts[i] = found;
    }
}

/* sparse/conv.cpp:46-100 (stride 1): per sample and output channel the
 * potentials start at 0; the events of each step, in list order, add
 * kernel[c][z][ky][kx] (one F32 add each) to every output position they
 * reach; after a step with events, every potential > threshold spikes
 * (output list order: channel, row, column) and becomes -FLT_MAX.  The
 * output is returned as its first-spike-time map [n][cout][ho][wo]. */
void orc_sparse_conv_times(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, int64_t h,
                           int64_t w, const float *kernel, int64_t cout, int64_t kh, int64_t kw, float threshold,
                           int64_t pt, int64_t pl, int64_t pb, int64_t pr, uint16_t *times)
{
    const int64_t ho = h + pt + pb - kh + 1, wo = w + pl + pr - kw + 1;
    float *pot = (float *)malloc(sizeof(float) * (size_t)(ho * wo));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t c = 0; c < cout; ++c)
        {
            uint16_t *tm = times + (i * cout + c) * ho * wo;
            for (int64_t p = 0; p < ho * wo; ++p) pot[p] = 0.0f, tm[p] = (uint16_t)t;
            for (int64_t j = 0; j < t; ++j)
            {
                const int64_t e0 = off[i * t + j], e1 = off[i * t + j + 1];
                for (int64_t e = e0; e < e1; ++e)
                {
                    const int64_t z = EZ(ev[e]), y = EY(ev[e]) + pt, x = EX(ev[e]) + pl;
                    for (int64_t oy = 0; oy < ho; ++oy)
                        for (int64_t ky = 0; ky < kh; ++ky)
                            for (int64_t kx = 0; kx < kw; ++kx)
                            {
                                const int64_t ox = x - kx;
                                if (oy + ky != y || ox < 0 || ox >= wo) continue;
                                pot[oy * wo + ox] += kernel[((c * cin + z) * kh + ky) * kw + kx];
                            }
                }
                if (e1 > e0)
                    for (int64_t p = 0; p < ho * wo; ++p)
                        if (pot[p] > threshold) tm[p] = (uint16_t)j, pot[p] = -FLT_MAX;
            }
        }
    free(pot);
}

int64_t orc_sparse_conv(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, int64_t h,
                        int64_t w, const float *kernel, int64_t cout, int64_t kh, int64_t kw, float threshold,
                        int64_t pt, int64_t pl, int64_t pb, int64_t pr, int64_t *oof, int32_t *oev, int64_t cap)
{
    const int64_t ho = h + pt + pb - kh + 1, wo = w + pl + pr - kw + 1;
    uint16_t *times = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n * cout * ho * wo));
    orc_sparse_conv_times(off, ev, n, t, cin, h, w, kernel, cout, kh, kw, threshold, pt, pl, pb, pr, times);
    const int64_t r = lists_from_times(times, n, t, cout * ho * wo, ho, wo, oof, oev, cap);
    free(times);
    return r;
}

/* sparse/stdp.cpp:57-136: for each winner (samples, then list order) and
 * each weight (z, ky, kx) of kernel[winner.z], the bit is "the last step
 * with an event at input (z, y + ky, x + kx) of the padded sample is <=
 * winner.t" (no event: t, so false); the update is the reference's F32
 * formula, its F64 terms included:
 *     mult = (F32)((k - lower) * (upper - k))         (1 without stabilize)
 *     k = (F32)fma(rate, mult, k)                     (old + rate * mult, contracted)
 *     k = (F32)max(lower, min(k, upper))
 * Returns -1 when a winner's config index is out of range (updates before
 * it stay applied, as in the reference). */
int32_t orc_sparse_stdp(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, float *kernel,
                        int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl,
                        const orc_stdp_config *configs, int64_t nconfigs, const orc_winner *winners,
                        const int64_t *counts, int64_t maxcount)
{
    const int64_t ksz = cin * kh * kw;
    uint16_t *direct = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)ksz);
    (void)cout;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < counts[i]; ++k)
        {
            const orc_winner *wn = &winners[i * maxcount + k];
            if (wn->c < 0) continue;
            if (wn->c >= nconfigs)
            {
                free(direct);
                return -1;
            }
            for (int64_t q = 0; q < ksz; ++q) direct[q] = (uint16_t)t;
            for (int64_t j = 0; j < t; ++j)
                for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
                {
                    const int64_t z = EZ(ev[e]), y = EY(ev[e]) + pt, x = EX(ev[e]) + pl;
                    if (wn->y <= y && y <= wn->y + kh - 1 && wn->x <= x && x <= wn->x + kw - 1)
                        direct[(z * kh + (y - wn->y)) * kw + (x - wn->x)] = (uint16_t)j;
                }
            const orc_stdp_config *cf = &configs[wn->c];
            float *kz = kernel + wn->z * ksz;
            for (int64_t q = 0; q < ksz; ++q)
            {
                const float mult = cf->stabilize ? (float)(((double)kz[q] - cf->lower) * (cf->upper - (double)kz[q]))
                                                 : 1.0f;
                kz[q] = (float)fma(direct[q] <= wn->t ? cf->positive : cf->negative, (double)mult, (double)kz[q]);
                const double hi = (double)kz[q] < cf->upper ? (double)kz[q] : cf->upper;
                kz[q] = (float)(cf->lower < hi ? hi : cf->lower);
            }
        }
    free(direct);
    return 0;
}

/* Integer branch of fcwta, src/spyker/cpu/wta.cpp:66-85,87-126: value =
 * t - first nonzero step (0 if none); first maximum; stop at 0; the
 * [x - radius, x + radius] window is cleared. */
void orc_rank_fcwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                       orc_winner *winners, int64_t *counts)
{
    uint16_t *temp = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)f);
    for (int64_t s = 0; s < n; ++s)
    {
        const uint8_t *xs = x + s * t * f;
        memset(temp, 0, sizeof(uint16_t) * (size_t)f);
        for (int64_t i = t - 1; i >= 0; --i)
            for (int64_t j = 0; j < f; ++j)
                if (xs[i * f + j]) temp[j] = (uint16_t)(t - i);
        int64_t k = 0;
        for (; k < count; ++k)
        {
            int64_t best = 0;
            for (int64_t j = 1; j < f; ++j)
                if (temp[best] < temp[j]) best = j;
            if (temp[best] == 0) break;
            winners[s * count + k] = (orc_winner){0, 0, 0, 0, best};
            const int64_t a = best - radius > 0 ? best - radius : 0, b = best + radius + 1 < f ? best + radius + 1 : f;
            for (int64_t j = a; j < b; ++j) temp[j] = 0;
        }
        counts[s] = k;
    }
    free(temp);
}

/* ------------------------------------------------------------------------ */
/* readout: first-spike times and back (src/spyker/cpu/time.cpp)           */
/* ------------------------------------------------------------------------ */

/* rank gather, time.cpp:40-64: out[i][j] = the first step t with
 * in[i][t][j] > threshold, T if none (threshold cast to the input type). */
void orc_rank_gather_u8(const uint8_t *in, int64_t n, int64_t t, int64_t len, uint8_t threshold, int32_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < len; ++j)
        {
            int32_t f = (int32_t)t;
            for (int64_t s = t - 1; s >= 0; --s)
                if (in[(i * t + s) * len + j] > threshold) f = (int32_t)s;
            out[i * len + j] = f;
        }
}

void orc_rank_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, int32_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < len; ++j)
        {
            int32_t f = (int32_t)t;
            for (int64_t s = t - 1; s >= 0; --s)
                if (in[(i * t + s) * len + j] > threshold) f = (int32_t)s;
            out[i * len + j] = f;
        }
}

/* rank scatter, time.cpp:66-84: out[i][s][j] = (s >= in[i][j]), the step
 * index an unsigned 32-bit value compared in the input's type. */
void orc_rank_scatter_f32(const float *in, int64_t n, int64_t len, int64_t t, uint8_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t s = 0; s < t; ++s)
            for (int64_t j = 0; j < len; ++j) out[(i * t + s) * len + j] = (float)(uint32_t)s >= in[i * len + j];
}

void orc_rank_scatter_u8(const uint8_t *in, int64_t n, int64_t len, int64_t t, uint8_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t s = 0; s < t; ++s)
            for (int64_t j = 0; j < len; ++j) out[(i * t + s) * len + j] = (uint32_t)s >= (uint32_t)in[i * len + j];
}

/* rate gather, time.cpp:86-92: the last step, copied. */
void orc_rate_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float *out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < len; ++j) out[i * len + j] = in[(i * t + t - 1) * len + j];
}
