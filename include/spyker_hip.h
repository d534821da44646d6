/*
 * spyker_hip.h — C ABI of the MI355X (gfx950) device backend for Spyker's
 * dense STDP layer path.
 *
 * Every entry point replaces one device-side function of the reference's
 * backend seam, Core::cuda_<op> (declared through the Declare macro,
 * src/spyker/impl.h:40-43), and is cited beside its declaration.  The
 * arguments are those of the seam flattened to plain pointers and sizes:
 *
 *   - tensors are device pointers to dense row-major buffers; a DynN view
 *     of the reference (src/spyker/base.h:363-491) becomes (ptr, type, dims);
 *   - Scalar arguments (src/spyker/utils.h:424-474) become double, converted
 *     to the tensor's element type inside, as Scalar::to does;
 *   - Len2/Len4 become explicit integers; pads are in the reference's order
 *     (top, left, bottom, right) = Len4{t, z, y, x};
 *   - Winners (a host std::vector, src/spyker/utils.h:502) become a device
 *     array of spk_winner records [n][count] plus a device count per sample,
 *     so a layer step never leaves the device;
 *   - every call takes the hipStream_t it is ordered on (NULL = the null
 *     stream) and is asynchronous unless stated.
 *
 * Status: 0 on success; otherwise a negative code and a message readable
 * with spk_last_error() (thread-local).  The C++ layer above turns a nonzero
 * status into std::runtime_error in the reference's message format
 * (src/spyker/utils.h:50-66).
 */
#ifndef SPYKER_HIP_H
#define SPYKER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types, in the order of Spyker::Type (src/spyker/utils.h:96-117). */
enum spk_type
{
    SPK_I8 = 0,
    SPK_I16 = 1,
    SPK_I32 = 2,
    SPK_I64 = 3,
    SPK_U8 = 4,
    SPK_U16 = 5,
    SPK_U32 = 6,
    SPK_U64 = 7,
    SPK_F16 = 8,
    SPK_F32 = 9,
    SPK_F64 = 10
};

enum spk_status
{
    SPK_OK = 0,
    SPK_ERR_ARGUMENT = -1, /* shape, type or value check failed */
    SPK_ERR_UNSUPPORTED = -2, /* element type not implemented on this device */
    SPK_ERR_RUNTIME = -3, /* HIP runtime error */
    SPK_ERR_CONFIG = -4 /* STDP configuration index out of range */
};

/* Winner: src/spyker/utils.h:483-499 (c = STDP config index, t = time,
 * z = channel, y/x = position). */
typedef struct spk_winner
{
    int64_t c, t, z, y, x;
} spk_winner;

/* STDPConfig: src/spyker/utils.h:505-524. */
typedef struct spk_stdp_config
{
    double positive, negative;
    int32_t stabilize;
    double lower, upper;
} spk_stdp_config;

/* Maximum number of STDP configurations one call accepts. */
#define SPK_MAX_CONFIGS 32

/* -- runtime ------------------------------------------------------------- */

/* Message of the last failing call on this thread. */
const char *spk_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int32_t spk_version(void);
/* Core::cuda_device_count / cuda_set_device / cuda_current_device / cuda_sync
 * (src/spyker/impl.h:75-80), and cuda_memory_total / cuda_memory_free
 * (impl.h:88-90) as one query of the current device. */
int32_t spk_device_count(int32_t *count);
int32_t spk_set_device(int32_t device);
int32_t spk_current_device(int32_t *device);
int32_t spk_memory_info(int64_t *free_bytes, int64_t *total_bytes);
int32_t spk_synchronize(void *stream);

/* Caching allocator, Core::cuda_alloc/dealloc (src/spyker/cuda/memory.cu:139-155)
 * and its queries cuda_memory_taken/used, cuda_cache_clear
 * (src/spyker/cuda/memory.cu:157-198). */
int32_t spk_alloc(void **ptr, int64_t bytes);
int32_t spk_dealloc(void *ptr);
int32_t spk_memory_taken(int64_t *bytes);
int32_t spk_memory_used(int64_t *bytes);
int32_t spk_cache_clear(void);
/* Core::copy on the device (src/spyker/impl.h:154-157, cuda_copy): copies
 * `bytes` between any two host or device buffers, ordered on `stream`; a copy
 * to or from pageable host memory returns when it is done. */
int32_t spk_copy(const void *src, void *dst, int64_t bytes, void *stream);

/* -- elementwise --------------------------------------------------------- */

/* Core::cuda_threshold(Dyn1 input, Scalar threshold, Scalar value), impl.h:240:
 * in place, x <= threshold ? value : x.  Types: F32, F16, F64. */
int32_t spk_threshold(void *x, int32_t type, int64_t n, double threshold, double value, void *stream);

/* Core::cuda_rank_fire(Dyn1 input, Dyn1 output, Scalar threshold), impl.h:256:
 * out = (in > threshold).  in: F32/F16/F64; out: U8 or F32. */
int32_t spk_rank_fire(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, double threshold,
                      void *stream);

/* Core::cuda_cast(Size size, Dyn input, Dyn output), src/spyker/impl.h:160:
 * out[i] = (out type) in[i] for any two of the 11 element types, as
 * cpu_cast's std::copy converts (src/spyker/cpu/memory.cpp:54-62); used for
 * Tensor::to(type).  Float to integer truncates. */
int32_t spk_cast(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, void *stream);

/* Core::cuda_fill(Size size, Dyn input, Scalar value), src/spyker/impl.h:165:
 * x[i] = value converted to the element type (cpu/memory.cpp:64-70). */
int32_t spk_fill(void *x, int32_t type, int64_t n, double value, void *stream);

/* Core::cuda_pad(Dyn3 input, Dyn3 output, Len4 pad, Scalar value), impl.h:236:
 * planes of [h][w] into [h+pt+pb][w+pl+pr] filled with value.  Types: U8, F16, F32, F64. */
int32_t spk_pad(const void *in, void *out, int32_t type, int64_t planes, int64_t h, int64_t w, int64_t pt, int64_t pl,
                int64_t pb, int64_t pr, double value, void *stream);

/* -- pooling, inhibition, winner-take-all --------------------------------- */

/* Core::cuda_rank_pool(Dyn4 input, Dyn4 output, Len2 kernel, Len2 stride,
 * Len4 pad), impl.h:268: per plane, zero pad then window max seeded with the
 * reference's Limits<T>::min() (the lowest value; for F16 the smallest
 * positive normal, base.h:761-774).  Types: U8, F16, F32, F64. */
int32_t spk_rank_pool(const void *in, void *out, int32_t type, int64_t planes, int64_t h, int64_t w, int64_t kh,
                      int64_t kw, int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream);

/* Core::cuda_rank_inhibit(Dyn4 input, Scalar threshold), impl.h:272:
 * x is [n][t][c][hw], in place, compared and scaled in its own type with the
 * threshold converted to it (cpu/inhibit.cpp:41-121).  Types: F16, F32, F64. */
int32_t spk_rank_inhibit(void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t hw, double threshold,
                         void *stream);

/* Core::cuda_rank_convwta(Dyn5 input, Len2 radius, Size count, Scalar
 * threshold), impl.h:280: x is [n][t][c][h][w]; winners (device) receives
 * n*count records, counts (device) n entries.  Types: F16, F32, F64 (values
 * computed in that type, the reference's Limits<T>::min() for unfired
 * neurons: the smallest positive normal for F16, cpu/wta.cpp:44-64 and
 * base.h:761-774), U8. */
int32_t spk_rank_convwta(const void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                         int64_t ry, int64_t rx, int64_t count, double threshold, spk_winner *winners,
                         int64_t *counts, void *stream);

/* Core::cuda_rank_fcwta(Dyn3 input, Size radius, Size count, Scalar
 * threshold), impl.h:276: x is [n][t][f].  Types: F16, F32, F64 (as
 * spk_rank_convwta), and U8 (the reference's integer branch, value = t -
 * first spike step, cpu/wta.cpp:66-85). */
int32_t spk_rank_fcwta(const void *x, int32_t type, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                       double threshold, spk_winner *winners, int64_t *counts, void *stream);

/* -- contractions and filters --------------------------------------------- */

/* Core::cuda_conv(Dyn4 input, Dyn4 kernel, Dyn4 output, Len2 stride, Len4
 * pad), impl.h:224: cross-correlation, zero pad.  in [n][cin][h][w] is
 * converted to the kernel's type first and out has the kernel's type
 * (opers.cpp:408-421): kernel F32 (in U8, F16, F32 or F64; out F32, or F64
 * for the high-precision check path), F16 (in U8, F16 or F32; out F16) or F64
 * (in U8, F32 or F64; out F64), [cout][cin][kh][kw].  Accumulation is one
 * fused multiply-add per tap in (cin, ky, kx) order starting from +0, in FP32
 * for F32 and F16 kernels (an F16 output is that sum rounded once) and in
 * FP64 for F64 kernels (and the check path). */
int32_t spk_conv(const void *in, int32_t in_type, const void *kernel, int32_t kernel_type, void *out,
                 int32_t out_type, int64_t n, int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw,
                 int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream);

/* Core::cuda_fc(Dyn3 input, Dyn2 kernel, Dyn3 output), impl.h:228:
 * out[r][o] = sum_f in[r][f]*k[o][f], one fused multiply-add per f in
 * order.  in: U8 or F32; kernel/out F32. */
int32_t spk_fc(const void *in, int32_t in_type, const void *kernel, void *out, int64_t rows, int64_t fin,
               int64_t fout, void *stream);

/* Core::cuda_dog(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad), impl.h:200:
 * in F32 [planes][h][w], kernel F32 [2][count][k][k], out F32
 * [planes][count][ho][wo] = max(g1 - g2, 0). */
int32_t spk_dog(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream);

/* Core::cuda_log(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad), impl.h:208:
 * in F32 [planes][h][w], kernel F32 [2][count][k][k], out F32
 * [planes][2][count][ho][wo] with d = g1 - g2: [0] = max(d, 0), [1] = max(-d, 0). */
int32_t spk_log(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream);

/* Core::cuda_gabor(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad), impl.h:204:
 * in F32 [planes][h][w], kernel F32 [count][k][k], out F32 [planes][count][ho][wo]. */
int32_t spk_gabor(const void *in, const void *kernel, void *out, int64_t planes, int64_t h, int64_t w, int64_t count,
                  int64_t k, int64_t pt, int64_t pl, int64_t pb, int64_t pr, void *stream);

/* -- whitening ------------------------------------------------------------ */

/* Core::zca_fit(Dyn2 input, Dyn1 mean, Dyn2 trans, Scalar epsilon, bool
 * transform), impl.h:212 (the reference runs it on the host, cpu/zca.cpp:66-105):
 * in F32 [rows][cols] (whitened in place when transform != 0), mean F32
 * [cols], trans F32 [cols][cols] = U diag(1 / sqrt(max(0, l) + eps)) U^T of
 * the covariance (divisor rows - 1), eps taken as F32.  Computed in FP64 on
 * the device (rocBLAS/rocSOLVER, opened on first use); rows >= 2,
 * cols <= 16384.  Synchronizes `stream` (the eigensolver's status is read). */
int32_t spk_zca_fit(void *in, void *mean, void *trans, int64_t rows, int64_t cols, double epsilon, int32_t transform,
                    void *stream);

/* Core::zca_trans(Dyn2 input, Dyn1 mean, Dyn2 trans), impl.h:216
 * (cpu/zca.cpp:107-121): in F32 [rows][cols] <- (in - mean) . T in place, T
 * the symmetric matrix of trans's row-major upper triangle; FP64 products,
 * rounded once to F32.  cols <= 16384. */
int32_t spk_zca_trans(void *in, const void *mean, const void *trans, int64_t rows, int64_t cols, void *stream);

/* Core::zca_split(Dyn2 input, Dyn3 output), impl.h:220 (cpu/zca.cpp:141-156):
 * in [rows][cols] -> out [rows][2][cols] = max(x, 0), max(-x, 0) with C++'s
 * conversions of -x to the element type.  Types: all eleven. */
int32_t spk_zca_split(const void *in, void *out, int32_t type, int64_t rows, int64_t cols, void *stream);

/* -- coding --------------------------------------------------------------- */

/* Core::cuda_rank_code(Dyn2 input, Dyn3 output, bool sort), impl.h:248:
 * in F32 [n][len] -> out U8 [n][time][len], cumulative spikes.  sort=1 ranks
 * by value (descending, ties by ascending index); sort=0 is latency coding. */
int32_t spk_rank_code(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                      int64_t time, int32_t sort, void *stream);

/* Core::cuda_rate_code(Dyn2 input, Dyn3 output, bool sort), impl.h:295, run
 * as the reference's CPU algorithm (cpu/code.cpp:134-243): in F32 [n][len] ->
 * out U8 or F32 [n][time][len], the spike count of each element so far.
 * Element j's rank-code time r (latency, or sorted when sort=1) selects row
 * min(r, time) of the Poisson table (spk_poisson_table with time + 1 rows);
 * sample i draws from std::mt19937(rnd + i), rnd one draw of the generator
 * that spk_random_seed seeds, taken when the call is made.  len <= 16384,
 * time <= 4096, n < 65536.  The first call for a `time` builds and uploads
 * its table (synchronous); later calls are asynchronous. */
int32_t spk_rate_code(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                      int64_t time, int32_t sort, void *stream);

/* Core::poisson_create(Size time, Size bins), impl.h:190 (cpu/init.cpp:120-149),
 * on the host: table [rows][bins] U16, row i the intervals of mean i + 1.
 * Each row's contents are the reference's; its order is a fixed permutation
 * (the reference's std::rand shuffle differs from run to run). */
int32_t spk_poisson_table(uint16_t *table, int64_t rows, int64_t bins);

/* Core::cuda_poisson_clear, impl.h:126: frees the device tables that
 * spk_rate_code caches per (device, time); waits for each device first. */
int32_t spk_poisson_clear(void);

/* Core::cuda_rate_fire(Dyn3 input, Dyn3 output, Scalar threshold), impl.h:299
 * (cpu/fire.cpp:84-114): in [n][t][len] -> out [n][t][len] spike counts; per
 * element a spike counts when value + threshold < x, and value becomes x
 * (value starts at 0, threshold converted to the input type).  Types (in and
 * out): U8, U16, I32, F32. */
int32_t spk_rate_fire(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                      int64_t len, double threshold, void *stream);

/* -- learning ------------------------------------------------------------- */

/* Core::cuda_rank_convstdp(Dyn5 input, Dyn4 kernel, Dyn5 output, configs,
 * Winners, Len2 stride, Len4 pad), impl.h:284-289.  in/out U8 spikes
 * [n][t][cin][h][w] / [n][t][cout][ho][wo]; kernel F32, F16 or F64
 * [cout][cin][kh][kw] updated in place in (sample, winner) order, each update
 * in F64 rounded once to the kernel's type (cpu/stdp.cpp:121-137); configs on the host; winners
 * and counts on the device ([n][maxcount], [n]).  Winners with c < 0 are
 * skipped; any c >= nconfigs fails the call before any weight changes
 * (SPK_ERR_CONFIG). */
int32_t spk_rank_convstdp(const void *in, void *kernel, int32_t kernel_type, const void *out, int64_t n, int64_t t,
                          int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw, int64_t ho,
                          int64_t wo, const spk_stdp_config *configs, int64_t nconfigs, const spk_winner *winners,
                          const int64_t *counts, int64_t maxcount, int64_t sy, int64_t sx, int64_t pt, int64_t pl,
                          int64_t pb, int64_t pr, void *stream);

/* Core::cuda_rank_fcstdp(Dyn3 input, Dyn2 kernel, Dyn3 output, configs,
 * Winners), impl.h:290-294.  in [n][t][fin] U8, out [n][t][fout] U8,
 * kernel F32, F16 or F64 [fout][fin]. */
int32_t spk_rank_fcstdp(const void *in, void *kernel, int32_t kernel_type, const void *out, int64_t n, int64_t t,
                        int64_t fin, int64_t fout, const spk_stdp_config *configs, int64_t nconfigs,
                        const spk_winner *winners, const int64_t *counts, int64_t maxcount, void *stream);

/* -- readout: first-spike times and back (cpu/time.cpp, cuda/time.cu) -------- */

/* Core::rank_gather (cuda_rank_gather, impl.h:260-262): out[i][j] = the first
 * step s with in[i][s][j] > threshold, t where none.  in [n][t][len], out
 * [n][len]; types U8, U16, I32, F32 (integer inputs take the threshold
 * truncated toward zero, clamped into their range). */
int32_t spk_rank_gather(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                        int64_t len, double threshold, void *stream);
/* Core::rank_scatter (cuda_rank_scatter, impl.h:264-266): out[i][s][j] =
 * (s >= in[i][j]) with s an unsigned 32-bit step index compared in the input's
 * type.  in [n][len], out [n][t][len]; types U8, U16, I32, F32. */
int32_t spk_rank_scatter(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t len,
                         int64_t t, void *stream);
/* Core::rate_gather (cuda_rate_gather, impl.h:303-305): out[i][j] =
 * in[i][t-1][j], converted (float to integer truncated and clamped). */
int32_t spk_rate_gather(const void *in, int32_t in_type, void *out, int32_t out_type, int64_t n, int64_t t,
                        int64_t len, void *stream);

/* -- sparse event path (Core::sparse_*, impl.h:328-356; src/spyker/sparse/) ------ */

/* A sparse spike tensor of shape [n][t][c][h][w] holds, for every sample i
 * and step j, the list of spikes that are new at step j (base.h:493-528).
 * Device layout: offsets[n*t + 1] (int64) and events[total] (int32 packed
 * z << 16 | y << 8 | x, the reference's U8 Spridx); list (i, j) is
 * events[offsets[i*t + j] .. offsets[i*t + j + 1]).  c, h and w are at most
 * 255, t at most 4096 and n at most 65535.  A function that makes a sparse tensor allocates its offsets and
 * events with spk_alloc and fills every field; release it with
 * spk_sparse_release.  Making one reads its size back to the host, so those
 * calls synchronize their stream. */
typedef struct spk_sparse
{
    int64_t n, t, c, h, w;
    int64_t total;
    int64_t *offsets;
    int32_t *events;
} spk_sparse;

/* Frees what a producer allocated and clears the pointers. */
int32_t spk_sparse_release(spk_sparse *s);

/* SparseTensor(tensor, threshold) (sparse/opers.cpp:38-80): an element's
 * event is at the first step whose value is > threshold.  in: [n][t][c][h][w],
 * types U8, F32. */
int32_t spk_sparse_from_dense(const void *in, int32_t type, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                              double threshold, spk_sparse *out, void *stream);

/* Sparse::code (Core::sparse_code, impl.h:350; sparse/code.cpp:98-132): rank
 * coding of F32 [n][c][h][w] into `time` steps; sort = 0 is latency coding. */
int32_t spk_sparse_code(const void *in, int32_t type, int64_t n, int64_t c, int64_t h, int64_t w, int64_t time,
                        int32_t sort, spk_sparse *out, void *stream);

/* SparseTensor::dense (Core::sparse_convert, impl.h:340): cumulative U8
 * spikes [n][t][c][h][w]. */
int32_t spk_sparse_to_dense(const spk_sparse *in, void *out, void *stream);

/* Sparse::gather (Core::sparse_gather, impl.h:348): first spike step of
 * every element, t where none, [n][c][h][w].  Types: U8, U16, I32, F32. */
int32_t spk_sparse_gather(const spk_sparse *in, void *out, int32_t type, void *stream);

/* Sparse::pad (Core::sparse_pad, impl.h:346): every event moves by (pt, pl). */
int32_t spk_sparse_pad(const spk_sparse *in, int64_t pt, int64_t pl, int64_t pb, int64_t pr, spk_sparse *out,
                       void *stream);

/* Sparse::pool (Core::sparse_pool, impl.h:344; sparse/pool.cpp): stride must
 * equal the kernel; the first event reaching a pooled cell is kept, in the
 * input's list order. */
int32_t spk_sparse_pool(const spk_sparse *in, int64_t kh, int64_t kw, int64_t sh, int64_t sw, int64_t pt, int64_t pl,
                        int64_t pb, int64_t pr, spk_sparse *out, void *stream);

/* Sparse::inhibit (Core::sparse_inhibit, impl.h:352): a pixel keeps the
 * channel of its first event. */
int32_t spk_sparse_inhibit(const spk_sparse *in, spk_sparse *out, void *stream);

/* Sparse::convwta (Core::sparse_convwta, impl.h:354; sparse/wta.cpp): first
 * come in step and list order, channel and box exclusion; winners/counts on
 * the device as for spk_rank_convwta; count at most 3072. */
int32_t spk_sparse_convwta(const spk_sparse *in, int64_t ry, int64_t rx, int64_t count, spk_winner *winners,
                           int64_t *counts, void *stream);

/* fcwta on spike lists.  No single reference call: it returns the winners
 * that Core::rank_fcwta's integer branch (cpu/wta.cpp:66-126) picks on
 * SparseTensor::dense() (utils.cpp:251-257) viewed as [n][t][c*h*w], without
 * materialising the dense tensor: candidates in (step, flat index) order,
 * flat = (z*h + y)*w + x, each a winner unless within `radius` of an earlier
 * one in flat index.  Each winner is returned as {c=0, t=step, z, y, x}
 * (the dense call returns {0, 0, 0, 0, flat}), so the list feeds
 * spk_sparse_convstdp directly.  count at most 8192. */
int32_t spk_sparse_fcwta(const spk_sparse *in, int64_t radius, int64_t count, spk_winner *winners, int64_t *counts,
                         void *stream);

/* Conv::operator()(SparseTensor, threshold) (Core::sparse_conv, impl.h:342;
 * sparse/conv.cpp): stride 1; per sample and output neuron the F32
 * potential adds kernel[c][z][ky][kx] once per input event in list order,
 * spikes the first time a step with events leaves it > threshold, and then
 * never again.  kernel: F32 [cout][c][kh][kw] (the reference reads any
 * kernel as F32), c*kh*kw at most 416, t at most 1024. */
int32_t spk_sparse_conv(const spk_sparse *in, const void *kernel, int32_t kernel_type, int64_t cout, int64_t kh,
                        int64_t kw, double threshold, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb,
                        int64_t pr, spk_sparse *out, void *stream);

/* Conv::stdp(SparseTensor, Winners) (Core::sparse_stdp, impl.h:356;
 * sparse/stdp.cpp): for each winner in (sample, rank) order and each weight
 * (z, ky, kx) of kernel[winner.z]: potentiate iff the last event at input
 * (z, winner.y + ky - pt, winner.x + kx - pl) is at a step <= winner.t; the
 * update is the sparse path's F32 rule.  Winners with c < 0 are skipped; a
 * c >= nconfigs fails the call before any weight changes. */
int32_t spk_sparse_convstdp(const spk_sparse *in, void *kernel, int32_t kernel_type, int64_t cout, int64_t kh,
                            int64_t kw, const spk_stdp_config *configs, int64_t nconfigs, const spk_winner *winners,
                            const int64_t *counts, int64_t maxcount, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                            void *stream);

/* -- fused layer step (FP16 MFMA) ------------------------------------------ */

/* One training step of a conv layer as the reference runs it in separate
 * calls (examples/kheradpisheh_mnist_dense.py:62-64):
 *
 *     pot = inhibit(threshold(conv(in), threshold))    opers.cpp:408,472,639
 *     winners = convwta(pot, (ry, rx), count)          opers.cpp:656
 *     out = fire(pot)                                  opers.cpp:522
 *     Conv::stdp(in, winners, out)                     opers.cpp:873
 *     pooled = pool(out, (pkh, pkw), (psh, psw))       opers.cpp:600
 *
 * with the inhibit/convwta/fire thresholds at their default 0, stride 1 and
 * threshold >= 0.  The conv runs on FP16 MFMA (weights = the FP32 masters
 * rounded to FP16, spikes exact, FP32 accumulation); the STDP update is
 * applied to the FP32 masters in `kernel` with the reference's order and
 * arithmetic.  in: U8 [n][t][cin][h][w]; kernel: F32 [cout][cin][kh][kw]
 * (updated in place when nconfigs > 0); out: U8 [n][t][cout][ho][wo] or NULL
 * (required for learning); pooled: U8 [n][t][cout][pho][pwo] or NULL;
 * winners/counts: device [n][count], [n], every slot written (zeros past a
 * sample's count, so they need no clearing); potentials: NULL, or F32
 * [n][t][cout][ho][wo] receiving the raw FP16-path potentials (a check
 * output).  Domain (SPK_ERR_UNSUPPORTED outside it, so a caller can run the
 * plain path instead): t <= 32, 2 <= cout <= 256, cin*kh*kw <= 512 after
 * padding cin to 8, threshold >= 0.  A sample whose input fits 160 KiB of
 * LDS whole (and ho*wo <= 8192) runs the whole-sample kernel; any other runs
 * the region kernel with a cached plan (ho*wo <= 65536, pool windows <= 16
 * pixels).  No counterpart in the reference as one call. */
int32_t spk_stdp_layer_fused(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, void *kernel,
                             int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                             double threshold, int64_t ry, int64_t rx, int64_t count,
                             const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled, int64_t pkh,
                             int64_t pkw, int64_t psh, int64_t psw, spk_winner *winners, int64_t *counts,
                             void *potentials, void *stream);

/* The arithmetic of spk_stdp_layer_fused's whole-sample kernel (no reference
 * counterpart): mode 0 (the default) runs the conv as int8 MFMAs on the spike
 * bytes with the weights in 16-bit fixed point (w rounded to a multiple of
 * 2^-e, e the largest with max|w| * 2^e <= 32639, chosen per call), whose
 * potentials are exact sums of the rounded weights; shapes whose window rows
 * need more than 8 K-steps of 64 bytes use FP16 operands instead.  Mode 1
 * uses FP16 operands (weights rounded to FP16) everywhere.  Both accumulate
 * exactly what they sum in their own format; samples too large for the LDS
 * (the region kernel) always use FP16.  Process-wide. */
int32_t spk_set_fused_arith(int32_t mode);

/* The whole-sample kernel's output channels a wave, its tile shape (the
 * reference picks its conv algorithm per shape at run time,
 * cuda/conv.cu:93-104,161-178; this is the fused step's counterpart): 32
 * (two waves a SIMD) or 64 (one wave a SIMD, half the operand reads a
 * multiply).  With tuning on (mode 1, the default) the first call of each
 * shape (T, Cin, H, W, Cout, Kh, Kw, pad, arithmetic) times both on its first
 * min(n, 2 x CUs) samples and keeps the faster for later calls; the two give
 * the same bytes.  No tuning under a stream capture, in the allocation-free
 * entry point or with the potentials output (32 then, or the kept choice).
 * Process-wide. */
int32_t spk_set_fused_tuning(int32_t mode);

/* Forces 32 or 64 channels a wave for every later call (0: the tuned choice,
 * else 32).  Process-wide. */
int32_t spk_set_fused_channels(int32_t channels);

/* The tuned channels a wave for a shape under the current arithmetic, or 0
 * where no call of that shape has been tuned. */
int32_t spk_fused_tuned_channels(int64_t t, int64_t cin, int64_t h, int64_t w, int64_t cout, int64_t kh, int64_t kw,
                                 int64_t pt, int64_t pl, int64_t pb, int64_t pr, int32_t *channels);

/* spk_stdp_layer_fused for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and results
 * as spk_stdp_layer_fused; the caller still passes every input and output.
 * The step's own buffers come from a 16 MiB device array that the library
 * holds on each GPU it runs on until the process ends, so the call makes no
 * allocation.  It runs only the whole-sample path: a call whose buffers need
 * more than 16 MiB (at the layer-2 shape, about 1700 samples) or whose
 * samples do not fit the LDS whole fails with SPK_ERR_ARGUMENT.  Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (from two
 * streams or threads) race. */
int32_t spk_stdp_layer_fused_noalloc(const void *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w,
                                     void *kernel, int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl,
                                     int64_t pb, int64_t pr, double threshold, int64_t ry, int64_t rx, int64_t count,
                                     const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled,
                                     int64_t pkh, int64_t pkw, int64_t psh, int64_t psw, spk_winner *winners,
                                     int64_t *counts, void *potentials, void *stream);

/* -- layer plans: batches of uneven instances ------------------------------- */

/* No single reference counterpart: the fused step above for a batch of n
 * instances of their own size and step count, the batch the reference would
 * run as n calls in order with one STDP pass over all winners afterwards
 * (the reference's batch order: every instance's forward pass on the weights
 * as they were, then the winners of instance 0, 1, ... in rank order).
 * shapes: host array [n][3] of (t, h, w); t <= 32, Ho*Wo <= 65536.  The plan
 * allocates its tables and work buffers once; spk_layer_plan_run then
 * allocates nothing and never waits on the host, so a stream capture of it
 * (hipGraph) can be replayed.  tune != 0: the first run times the region
 * heights that fit and keeps the fastest (run it once before capturing).
 * pkh = 0: no pooling. */
typedef struct spk_layer_plan spk_layer_plan;
int32_t spk_layer_plan_create(spk_layer_plan **plan, int64_t n, const int32_t *shapes, int64_t cin, int64_t cout,
                              int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t pb, int64_t pr,
                              double threshold, int64_t ry, int64_t rx, int64_t count, int64_t pkh, int64_t pkw,
                              int64_t psh, int64_t psw, int32_t tune);
/* Byte offsets of each instance in the packed buffers (host arrays [n], any
 * may be null): input [t][cin][h][w], spikes [t][cout][ho][wo], pooled
 * spikes [t][cout][pho][pwo]; totals[3] = their total sizes. */
int32_t spk_layer_plan_layout(const spk_layer_plan *plan, int64_t *in_off, int64_t *out_off, int64_t *pool_off,
                              int64_t *totals);
/* One step: packed U8 input, the FP32 master kernel [cout][cin][kh][kw]
 * (updated in place when nconfigs > 0; every winner uses configuration 0),
 * packed spikes and pooled spikes (either may be null), winners [n][count]
 * and counts [n] as for spk_rank_convwta. */
int32_t spk_layer_plan_run(spk_layer_plan *plan, const void *in, void *kernel, int32_t kernel_type,
                           const spk_stdp_config *configs, int64_t nconfigs, void *out, void *pooled,
                           spk_winner *winners, int64_t *counts, void *stream);
int32_t spk_layer_plan_destroy(spk_layer_plan *plan);
/* The plan's candidate tilings (no reference counterpart: the reference picks
 * cuDNN algorithms per shape at run time, src/spyker/cuda/conv.cu:161-178).
 * Candidate i has rows[i] output rows per region and covers the longest
 * instance's steps in chunks[i] passes of the FP16 windows.  *count gets the
 * number of candidates (at most capacity are written), *chosen the one runs
 * use, *writes the write mode: 1 = the region kernel writes the spikes (and
 * pooled spikes whose windows lie inside one group of its rows), 0 = separate
 * writer kernels.  A plan made with tune != 0 times every candidate in both
 * modes on its first run and keeps the fastest; otherwise it uses the first
 * candidate and a mode chosen from the shape.  spk_layer_plan_select fixes
 * candidate `index` and mode `writes` (-1 keeps the mode) for every later
 * run, for a caller that tuned once and restores its choice. */
int32_t spk_layer_plan_candidates(const spk_layer_plan *plan, int32_t *rows, int32_t *chunks, int32_t capacity,
                                  int32_t *count, int32_t *chosen, int32_t *writes);
int32_t spk_layer_plan_select(spk_layer_plan *plan, int32_t index, int32_t writes);


/* HIP-event timing of the main kernel of spk_stdp_layer_fused, recorded on
 * the launch stream around each launch while enabled (no reference
 * counterpart; the reference's unused Timer is src/spyker/base.h:80-95).
 * spk_kernel_timer(1) clears and starts, spk_kernel_timer(0) clears and
 * stops; spk_kernel_timer_read waits for the recorded launches and returns
 * their summed duration and count. */
int32_t spk_kernel_timer(int32_t enable);
int32_t spk_kernel_timer_read(double *total_ms, int64_t *launches);

/* -- host kernel generators (F64, host memory) ---------------------------- */

/* Core::random_seed (src/spyker/impl.cpp:43): reseeds the std::mt19937 that
 * spk_normal_kernel draws from. */
int32_t spk_random_seed(int64_t seed);
/* Core::normal_kernel (src/spyker/cpu/init.cpp:45-49): std::normal_distribution<double>. */
int32_t spk_normal_kernel(double *kernel, int64_t n, double mean, double std);
/* Core::gaussian_kernel (src/spyker/cpu/init.cpp:51-62): k x k, sum-normalised. */
int32_t spk_gaussian_kernel(double *kernel, int64_t k, double std);
/* Core::gabor_kernel (src/spyker/cpu/init.cpp:64-77): k x k, sum-normalised, theta in radians. */
int32_t spk_gabor_kernel(double *kernel, int64_t k, double sigma, double theta, double gamma, double lambda,
                         double psi);

#ifdef __cplusplus
}
#endif
#endif
