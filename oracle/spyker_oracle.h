/*
 * CPU oracle for the Spyker dense STDP layer path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (spyker_amd/) links,
 * loads or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function restates one function of the reference's CPU backend in
 * plain C; the file:line each one follows is given beside it.  Layouts are
 * row-major with time outside channel: [N][T][C][H][W].
 *
 * Pads are given in the reference's order (top, left, bottom, right), which
 * is the order of Len4{t,z,y,x} in src/spyker/cpu/pool.cpp:41-51.
 */
#ifndef SPYKER_ORACLE_H
#define SPYKER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    double positive, negative;
    int32_t stabilize;
    double lower, upper;
} orc_stdp_config;

/* Winner record: c, t, z, y, x as in src/spyker/utils.h:483-499. */
typedef struct
{
    int64_t c, t, z, y, x;
} orc_winner;

/* -- rank-order coding (src/spyker/cpu/code.cpp) ------------------------- */
void orc_rank_code_sort_f32(const float *in, int64_t n, int64_t len, int64_t time, uint8_t *out, int32_t threads);
void orc_rank_code_latency_f32(const float *in, int64_t n, int64_t len, int64_t time, uint8_t *out);
/* spike times only (U16), for tie analysis */
void orc_rank_times_sort_f32(const float *in, int64_t len, int64_t time, uint16_t *times);

/* -- elementwise (src/spyker/cpu/fire.cpp) ------------------------------- */
void orc_threshold_f32(float *x, int64_t n, float threshold, float value);
void orc_rank_fire_f32(const float *x, uint8_t *out, int64_t n, float threshold);

/* -- inhibit (src/spyker/cpu/inhibit.cpp:41-121) ------------------------- */
void orc_rank_inhibit_f32(float *x, int64_t n, int64_t t, int64_t c, int64_t hw, float threshold, int32_t threads);

/* -- winner-take-all (src/spyker/cpu/wta.cpp) ---------------------------- */
/* winners: n*count records; counts: n entries. */
void orc_rank_convwta_f32(const float *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                          int64_t rx, int64_t count, float threshold, orc_winner *winners, int64_t *counts,
                          int32_t threads);
void orc_rank_convwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                         int64_t rx, int64_t count, orc_winner *winners, int64_t *counts);
void orc_rank_fcwta_f32(const float *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                        float threshold, orc_winner *winners, int64_t *counts);

/* -- pooling / padding (src/spyker/cpu/pool.cpp:41-86) ------------------- */
void orc_pad_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t pt, int64_t pl,
                int64_t pb, int64_t pr, uint8_t value);
void orc_pad_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t pt, int64_t pl, int64_t pb,
                 int64_t pr, float value);
void orc_rank_pool_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t kh, int64_t kw,
                      int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr);
void orc_rank_pool_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t kh, int64_t kw,
                       int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr);

/* -- STDP (src/spyker/cpu/stdp.cpp) -------------------------------------- */
/* Returns 0 on success, -1 when a winner's config index is out of range. */
int32_t orc_rank_convstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w,
                              float *kernel, int64_t cout, int64_t kh, int64_t kw, const uint8_t *out, int64_t ho,
                              int64_t wo, const orc_stdp_config *configs, int64_t nconfigs,
                              const orc_winner *winners, const int64_t *counts, int64_t maxcount, int64_t sy,
                              int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr);
int32_t orc_rank_fcstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t fin, float *kernel, int64_t fout,
                            const uint8_t *out, const orc_stdp_config *configs, int64_t nconfigs,
                            const orc_winner *winners, const int64_t *counts, int64_t maxcount);
/* The scalar weight update alone, for exhaustive checks. */
float orc_stdp_update(float w, int32_t potentiate, const orc_stdp_config *cfg);

/* -- filters and contractions -------------------------------------------- */
/* Canonical FP32 cross-correlation: acc starts at +0 and takes fmaf over
 * (c, ky, kx) in that nesting order; taps that fall in the zero pad are
 * skipped.  in is U8 (is_u8=1) or F32. */
void orc_conv_f32(const void *in, int32_t is_u8, int64_t n, int64_t cin, int64_t h, int64_t w, const float *kernel,
                  int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                  int64_t pr, float *out, int32_t threads);
/* FP64 potentials (the reduced-precision path's reference). */
void orc_conv_f64(const void *in, int32_t is_u8, int64_t n, int64_t cin, int64_t h, int64_t w, const float *kernel,
                  int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                  int64_t pr, double *out, int32_t threads);
void orc_conv_f64k(const void *in, int32_t in_kind, int64_t n, int64_t cin, int64_t h, int64_t w,
                   const double *kernel, int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx, int64_t pt,
                   int64_t pl, int64_t pb, int64_t pr, double *out, int32_t threads);
void orc_fc_f32(const void *in, int32_t is_u8, int64_t rows, int64_t fin, const float *kernel, int64_t fout,
                float *out);
/* DoG: conv with [2*count][1][k][k], then max(g1-g2, 0) (src/spyker/cpu/dog.cpp:47-60). */
void orc_dog_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count, int64_t k,
                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads);
/* Gabor: plain conv with [count][1][k][k] (src/spyker/cpu/dog.cpp:94). */
void orc_gabor_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count,
                   int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads);

/* -- host kernel generators (src/spyker/cpu/init.cpp:38-77) --------------- */
void orc_log_f32(const float *in, int64_t planes, int64_t h, int64_t w, const float *kernel, int64_t count, int64_t k,
                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, float *out, int32_t threads);
void orc_mt_stream(uint32_t seed, int64_t n, uint32_t *out);
int32_t orc_rate_code_f32(const float *in, int64_t n, int64_t len, int64_t time, int32_t sort, uint64_t rnd,
                          const uint16_t *table, int64_t rows, int64_t bins, uint8_t *out);
void orc_rate_fire_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, uint8_t *out);
void orc_gaussian_kernel(double *kernel, int64_t k, double std);
void orc_gabor_kernel(double *kernel, int64_t k, double sigma, double theta, double gamma, double lambda, double psi);

/* -- sparse event path (src/spyker/sparse/) ------------------------------ */
/* A list set: off[n*t + 1] and events packed z << 16 | y << 8 | x; list
 * (i, j) is [off[i*t+j], off[i*t+j+1]).  Producers return the event total,
 * or -1 when more than `cap` events would be written. */
int64_t orc_sparse_code_f32(const float *in, int64_t n, int64_t c, int64_t h, int64_t w, int64_t t, int32_t sort,
                            int64_t *off, int32_t *ev, int64_t cap);
int64_t orc_sparse_from_dense_u8(const uint8_t *in, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                                 uint8_t threshold, int64_t *off, int32_t *ev, int64_t cap);
void orc_sparse_to_dense_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                            int64_t w, uint8_t *out);
void orc_sparse_gather_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                          int64_t w, uint8_t *out);
void orc_sparse_pad(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t pt, int64_t pl,
                    int32_t *out);
int64_t orc_sparse_pool(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t ho,
                        int64_t wo, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t *oof, int32_t *oev,
                        int64_t cap);
int64_t orc_sparse_inhibit(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                           int64_t *oof, int32_t *oev, int64_t cap);
void orc_sparse_convwta(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                        int64_t ry, int64_t rx, int64_t count, orc_winner *winners, int64_t *counts);
void orc_sparse_conv_times(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, int64_t h,
                           int64_t w, const float *kernel, int64_t cout, int64_t kh, int64_t kw, float threshold,
                           int64_t pt, int64_t pl, int64_t pb, int64_t pr, uint16_t *times);
int64_t orc_sparse_conv(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, int64_t h,
                        int64_t w, const float *kernel, int64_t cout, int64_t kh, int64_t kw, float threshold,
                        int64_t pt, int64_t pl, int64_t pb, int64_t pr, int64_t *oof, int32_t *oev, int64_t cap);
int32_t orc_sparse_stdp(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, float *kernel,
                        int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl,
                        const orc_stdp_config *configs, int64_t nconfigs, const orc_winner *winners,
                        const int64_t *counts, int64_t maxcount);
void orc_rank_fcwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                       orc_winner *winners, int64_t *counts);

/* readout (cpu/time.cpp) */
void orc_rank_gather_u8(const uint8_t *in, int64_t n, int64_t t, int64_t len, uint8_t threshold, int32_t *out);
void orc_rank_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, int32_t *out);
void orc_rank_scatter_f32(const float *in, int64_t n, int64_t len, int64_t t, uint8_t *out);
void orc_rank_scatter_u8(const uint8_t *in, int64_t n, int64_t len, int64_t t, uint8_t *out);
void orc_rate_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float *out);

#ifdef __cplusplus
}
#endif
#endif
