// C entry points over the reference's own CPU backend, compiled from the
// reference sources where they lie (see oracle/Makefile).  TEST
// INFRASTRUCTURE ONLY: it is the checker that pins the C oracle; the product
// never links or loads it.
//
// Each entry point wraps one Core::cpu_* function declared in
// src/spyker/impl.h:196-322 and has the signature of its orc_* twin in
// spyker_oracle.h, so tests can call both with the same arguments.

#include <spyker/impl.h>

#include <cstring>
#include <string>

using namespace Spyker;
using namespace Spyker::Core;

namespace
{
thread_local std::string last_error;

struct RefWinner
{
    int64_t c, t, z, y, x;
};
struct RefConfig
{
    double positive, negative;
    int32_t stabilize;
    double lower, upper;
};

template <typename F>
int guard(F f)
{
    try
    {
        f();
        return 0;
    }
    catch (const std::exception &e)
    {
        last_error = e.what();
        return -1;
    }
}

void put(const Winners &w, int64_t count, RefWinner *out, int64_t *counts)
{
    for (size_t s = 0; s < w.size(); ++s)
    {
        counts[s] = int64_t(w[s].size());
        for (size_t k = 0; k < w[s].size(); ++k)
            out[s * count + k] = {w[s][k].c, w[s][k].t, w[s][k].z, w[s][k].y, w[s][k].x};
    }
}

Winners get(const RefWinner *win, const int64_t *counts, int64_t n, int64_t maxcount)
{
    Winners w(n);
    for (int64_t s = 0; s < n; ++s)
        for (int64_t k = 0; k < counts[s]; ++k)
        {
            const RefWinner &r = win[s * maxcount + k];
            w[s].push_back(Winner{r.c, r.t, r.z, r.y, r.x});
        }
    return w;
}

std::vector<STDPConfig> configs(const RefConfig *c, int64_t n);

// list set (off, packed events) <-> the reference's Sparse5
Sparse5 to_sparse(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w)
{
    Sparse5 s = sparse_alloc(Len5(n, t, c, h, w));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < t; ++j)
            for (int64_t e = off[i * t + j]; e < off[i * t + j + 1]; ++e)
                s(i, j).add(Spridx(U8((ev[e] >> 16) & 0xFF), U8((ev[e] >> 8) & 0xFF), U8(ev[e] & 0xFF)));
    return s;
}

int64_t from_sparse(Sparse5 s, int64_t *off, int32_t *ev, int64_t cap)
{
    int64_t total = 0;
    off[0] = 0;
    for (int64_t i = 0; i < s.u; ++i)
        for (int64_t j = 0; j < s.t; ++j)
        {
            for (const Spridx &x : s(i, j))
            {
                if (total >= cap) return -1;
                ev[total++] = int32_t(uint32_t(x.z) << 16 | uint32_t(x.y) << 8 | uint32_t(x.x));
            }
            off[i * s.t + j + 1] = total;
        }
    return total;
}

std::vector<STDPConfig> configs(const RefConfig *c, int64_t n)
{
    std::vector<STDPConfig> out;
    for (int64_t i = 0; i < n; ++i)
        out.push_back(STDPConfig(c[i].positive, c[i].negative, c[i].stabilize != 0, c[i].lower, c[i].upper));
    return out;
}
}  // namespace

extern "C" {

const char *ref_last_error() { return last_error.c_str(); }

int ref_rank_code_f32(const float *in, int64_t n, int64_t len, int64_t time, uint8_t *out, int32_t sort)
{
    return guard([&] {
        cpu_rank_code(Dyn2((void *)in, Type::F32, n, len), Dyn3(out, Type::U8, n, time, len), sort != 0);
    });
}

int ref_threshold_f32(float *x, int64_t n, float threshold, float value)
{
    return guard([&] { cpu_threshold(Dyn1(x, Type::F32, n), Scalar(threshold), Scalar(value)); });
}

int ref_rank_fire_f32(const float *x, uint8_t *out, int64_t n, float threshold)
{
    return guard([&] { cpu_rank_fire(Dyn1((void *)x, Type::F32, n), Dyn1(out, Type::U8, n), Scalar(threshold)); });
}

int ref_rank_inhibit_f32(float *x, int64_t n, int64_t t, int64_t c, int64_t hw, float threshold)
{
    return guard([&] { cpu_rank_inhibit(Dyn4(x, Type::F32, n, t, c, hw), Scalar(threshold)); });
}

int ref_rank_convwta_f32(const float *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                         int64_t rx, int64_t count, float threshold, RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_convwta(Dyn5((void *)x, Type::F32, n, t, c, h, w), Len2(ry, rx), count, Scalar(threshold));
        put(r, count, winners, counts);
    });
}

int ref_rank_convwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w, int64_t ry,
                        int64_t rx, int64_t count, RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_convwta(Dyn5((void *)x, Type::U8, n, t, c, h, w), Len2(ry, rx), count, Scalar(U8(0)));
        put(r, count, winners, counts);
    });
}

int ref_rank_fcwta_f32(const float *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                       float threshold, RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_fcwta(Dyn3((void *)x, Type::F32, n, t, f), radius, count, Scalar(threshold));
        put(r, count, winners, counts);
    });
}

// cpu/pool.cpp:41-51 (Core::cpu_pad, impl.h:236): planes of [h][w] into
// [h+pt+pb][w+pl+pr], the border filled with value
int ref_pad_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t pt, int64_t pl,
               int64_t pb, int64_t pr, uint8_t value)
{
    return guard([&] {
        cpu_pad(Dyn3((void *)in, Type::U8, planes, h, w), Dyn3(out, Type::U8, planes, h + pt + pb, w + pl + pr),
                Len4(pt, pl, pb, pr), Scalar(value));
    });
}

int ref_pad_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t pt, int64_t pl, int64_t pb,
                int64_t pr, float value)
{
    return guard([&] {
        cpu_pad(Dyn3((void *)in, Type::F32, planes, h, w), Dyn3(out, Type::F32, planes, h + pt + pb, w + pl + pr),
                Len4(pt, pl, pb, pr), Scalar(value));
    });
}

int ref_rank_pool_u8(const uint8_t *in, int64_t planes, int64_t h, int64_t w, uint8_t *out, int64_t kh, int64_t kw,
                     int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    int64_t oh = (h + pt + pb - kh) / sh + 1, ow = (w + pl + pr - kw) / sw + 1;
    return guard([&] {
        cpu_rank_pool(Dyn4((void *)in, Type::U8, 1, planes, h, w), Dyn4(out, Type::U8, 1, planes, oh, ow),
                      Len2(kh, kw), Len2(sh, sw), Len4(pt, pl, pb, pr));
    });
}

int ref_rank_pool_f32(const float *in, int64_t planes, int64_t h, int64_t w, float *out, int64_t kh, int64_t kw,
                      int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    int64_t oh = (h + pt + pb - kh) / sh + 1, ow = (w + pl + pr - kw) / sw + 1;
    return guard([&] {
        cpu_rank_pool(Dyn4((void *)in, Type::F32, 1, planes, h, w), Dyn4(out, Type::F32, 1, planes, oh, ow),
                      Len2(kh, kw), Len2(sh, sw), Len4(pt, pl, pb, pr));
    });
}

int ref_rank_convstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, float *kernel,
                          int64_t cout, int64_t kh, int64_t kw, const uint8_t *out, int64_t ho, int64_t wo,
                          const RefConfig *cfg, int64_t ncfg, const RefWinner *winners, const int64_t *counts,
                          int64_t maxcount, int64_t sy, int64_t sx, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    return guard([&] {
        cpu_rank_convstdp(Dyn5((void *)in, Type::U8, n, t, cin, h, w), Dyn4(kernel, Type::F32, cout, cin, kh, kw),
                          Dyn5((void *)out, Type::U8, n, t, cout, ho, wo), configs(cfg, ncfg),
                          get(winners, counts, n, maxcount), Len2(sy, sx), Len4(pt, pl, pb, pr));
    });
}

int ref_rank_fcstdp_f32(const uint8_t *in, int64_t n, int64_t t, int64_t fin, float *kernel, int64_t fout,
                        const uint8_t *out, const RefConfig *cfg, int64_t ncfg, const RefWinner *winners,
                        const int64_t *counts, int64_t maxcount)
{
    return guard([&] {
        cpu_rank_fcstdp(Dyn3((void *)in, Type::U8, n, t, fin), Dyn2(kernel, Type::F32, fout, fin),
                        Dyn3((void *)out, Type::U8, n, t, fout), configs(cfg, ncfg),
                        get(winners, counts, n, maxcount));
    });
}

// The same entry points for any real type (Type ordinals F16 = 8, F32 = 9,
// F64 = 10, utils.h:96-117): the reference computes inhibit, wta, fire and
// the STDP update in the tensor's own type (F16 through half_float::half).
int ref_threshold_t(void *x, int32_t type, int64_t n, double threshold, double value)
{
    return guard([&] { cpu_threshold(Dyn1(x, Type(type), n), Scalar(threshold), Scalar(value)); });
}

int ref_rank_fire_t(const void *x, int32_t type, uint8_t *out, int64_t n, double threshold)
{
    return guard([&] { cpu_rank_fire(Dyn1((void *)x, Type(type), n), Dyn1(out, Type::U8, n), Scalar(threshold)); });
}

int ref_rank_inhibit_t(void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t hw, double threshold)
{
    return guard([&] { cpu_rank_inhibit(Dyn4(x, Type(type), n, t, c, hw), Scalar(threshold)); });
}

int ref_rank_convwta_t(const void *x, int32_t type, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                       int64_t ry, int64_t rx, int64_t count, double threshold, RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_convwta(Dyn5((void *)x, Type(type), n, t, c, h, w), Len2(ry, rx), count, Scalar(threshold));
        put(r, count, winners, counts);
    });
}

int ref_rank_fcwta_t(const void *x, int32_t type, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                     double threshold, RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_fcwta(Dyn3((void *)x, Type(type), n, t, f), radius, count, Scalar(threshold));
        put(r, count, winners, counts);
    });
}

int ref_rank_convstdp_t(const uint8_t *in, int64_t n, int64_t t, int64_t cin, int64_t h, int64_t w, void *kernel,
                        int32_t ktype, int64_t cout, int64_t kh, int64_t kw, const uint8_t *out, int64_t ho,
                        int64_t wo, const RefConfig *cfg, int64_t ncfg, const RefWinner *winners,
                        const int64_t *counts, int64_t maxcount, int64_t sy, int64_t sx, int64_t pt, int64_t pl,
                        int64_t pb, int64_t pr)
{
    return guard([&] {
        cpu_rank_convstdp(Dyn5((void *)in, Type::U8, n, t, cin, h, w), Dyn4(kernel, Type(ktype), cout, cin, kh, kw),
                          Dyn5((void *)out, Type::U8, n, t, cout, ho, wo), configs(cfg, ncfg),
                          get(winners, counts, n, maxcount), Len2(sy, sx), Len4(pt, pl, pb, pr));
    });
}

int ref_rank_pool_t(const void *in, int32_t type, int64_t planes, int64_t h, int64_t w, void *out, int64_t kh,
                    int64_t kw, int64_t sh, int64_t sw, int64_t pt, int64_t pl, int64_t pb, int64_t pr)
{
    int64_t oh = (h + pt + pb - kh) / sh + 1, ow = (w + pl + pr - kw) / sw + 1;
    return guard([&] {
        cpu_rank_pool(Dyn4((void *)in, Type(type), 1, planes, h, w), Dyn4(out, Type(type), 1, planes, oh, ow),
                      Len2(kh, kw), Len2(sh, sw), Len4(pt, pl, pb, pr));
    });
}

int ref_pad_t(const void *in, int32_t type, int64_t planes, int64_t h, int64_t w, void *out, int64_t pt, int64_t pl,
              int64_t pb, int64_t pr, double value)
{
    return guard([&] {
        cpu_pad(Dyn3((void *)in, Type(type), planes, h, w), Dyn3(out, Type(type), planes, h + pt + pb, w + pl + pr),
                Len4(pt, pl, pb, pr), Scalar(value));
    });
}

int ref_gaussian_kernel(double *kernel, int64_t k, double std)
{
    return guard([&] { gaussian_kernel(Vec2<F64>(kernel, k, k), std); });
}

int ref_gabor_kernel(double *kernel, int64_t k, double sigma, double theta, double gamma, double lambda, double psi)
{
    return guard([&] { gabor_kernel(Vec2<F64>(kernel, k, k), sigma, theta, gamma, lambda, psi); });
}

int ref_normal_kernel(double *kernel, int64_t n, double mean, double std, int64_t seed)
{
    return guard([&] {
        random_seed(seed);
        normal_kernel(Vec1<F64>(kernel, n), mean, std);
    });
}

// -- sparse event path (Core::sparse_*, src/spyker/impl.h:328-356) ----------

int64_t ref_sparse_code_f32(const float *in, int64_t n, int64_t c, int64_t h, int64_t w, int64_t t, int32_t sort,
                            int64_t *off, int32_t *ev, int64_t cap)
{
    int64_t r = -2;
    guard([&] {
        Sparse5 s = sparse_alloc(Len5(n, t, c, h, w));
        sparse_code(Dyn4((void *)in, Type::F32, n, c, h, w), s, sort != 0);
        r = from_sparse(s, off, ev, cap);
        sparse_dealloc(s);
    });
    return r;
}

int64_t ref_sparse_from_dense_u8(const uint8_t *in, int64_t n, int64_t t, int64_t c, int64_t h, int64_t w,
                                 uint8_t threshold, int64_t *off, int32_t *ev, int64_t cap)
{
    int64_t r = -2;
    guard([&] {
        Sparse5 s = sparse_alloc(Len5(n, t, c, h, w));
        sparse_convert(Dyn5((void *)in, Type::U8, n, t, c, h, w), s, Scalar(threshold));
        r = from_sparse(s, off, ev, cap);
        sparse_dealloc(s);
    });
    return r;
}

void ref_sparse_to_dense_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                            int64_t w, uint8_t *out)
{
    guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, c, h, w);
        sparse_convert(s, Dyn5(out, Type::U8, n, t, c, h, w));
        sparse_dealloc(s);
    });
}

void ref_sparse_gather_u8(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t h,
                          int64_t w, uint8_t *out)
{
    guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, c, h, w);
        sparse_gather(s, Dyn4(out, Type::U8, n, c, h, w));
        sparse_dealloc(s);
    });
}

void ref_sparse_pad(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t pt, int64_t pl,
                    int32_t *out)
{
    guard([&] {
        // shape only bounds the coordinates; 255 is the Spridx limit
        Sparse5 s = to_sparse(off, ev, n, t, 255, 255, 255), o = sparse_alloc(Len5(n, t, 255, 255, 255));
        sparse_pad(s, o, Len4(pt, pl, 0, 0));
        std::vector<int64_t> oo(size_t(n * t + 1));
        from_sparse(o, oo.data(), out, off[n * t]);
        sparse_dealloc(s), sparse_dealloc(o);
    });
}

int64_t ref_sparse_pool(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t c, int64_t ho,
                        int64_t wo, int64_t kh, int64_t kw, int64_t pt, int64_t pl, int64_t *oof, int32_t *oev,
                        int64_t cap)
{
    int64_t r = -2;
    guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, c, 255, 255), o = sparse_alloc(Len5(n, t, c, ho, wo));
        sparse_pool(s, o, Len2(kh, kw), Len2(kh, kw), Len4(pt, pl, 0, 0));
        r = from_sparse(o, oof, oev, cap);
        sparse_dealloc(s), sparse_dealloc(o);
    });
    return r;
}

int64_t ref_sparse_inhibit(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                           int64_t *oof, int32_t *oev, int64_t cap)
{
    int64_t r = -2;
    guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, 255, h, w), o = sparse_alloc(Len5(n, t, 255, h, w));
        sparse_inhibit(s, o);
        r = from_sparse(o, oof, oev, cap);
        sparse_dealloc(s), sparse_dealloc(o);
    });
    return r;
}

void ref_sparse_convwta(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t h, int64_t w,
                        int64_t ry, int64_t rx, int64_t count, RefWinner *winners, int64_t *counts)
{
    guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, 255, h, w);
        put(sparse_convwta(s, Len2(ry, rx), count), count, winners, counts);
        sparse_dealloc(s);
    });
}

int64_t ref_sparse_conv(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, int64_t h,
                        int64_t w, const float *kernel, int64_t cout, int64_t kh, int64_t kw, float threshold,
                        int64_t pt, int64_t pl, int64_t pb, int64_t pr, int64_t *oof, int32_t *oev, int64_t cap)
{
    int64_t r = -2;
    guard([&] {
        const int64_t ho = h + pt + pb - kh + 1, wo = w + pl + pr - kw + 1;
        Sparse5 s = to_sparse(off, ev, n, t, cin, h, w), o = sparse_alloc(Len5(n, t, cout, ho, wo));
        sparse_conv(s, Dyn4((void *)kernel, Type::F32, cout, cin, kh, kw), Scalar(threshold), o, Len2(1, 1),
                    Len4(pt, pl, pb, pr));
        r = from_sparse(o, oof, oev, cap);
        sparse_dealloc(s), sparse_dealloc(o);
    });
    return r;
}

int32_t ref_sparse_stdp(const int64_t *off, const int32_t *ev, int64_t n, int64_t t, int64_t cin, float *kernel,
                        int64_t cout, int64_t kh, int64_t kw, int64_t pt, int64_t pl, const RefConfig *cfg,
                        int64_t ncfg, const RefWinner *winners, const int64_t *counts, int64_t maxcount)
{
    return guard([&] {
        Sparse5 s = to_sparse(off, ev, n, t, cin, 255, 255);
        try
        {
            sparse_stdp(s, Dyn4(kernel, Type::F32, cout, cin, kh, kw), configs(cfg, ncfg),
                        get(winners, counts, n, maxcount), Len4(pt, pl, 0, 0));
        }
        catch (...)
        {
            sparse_dealloc(s);
            throw;
        }
        sparse_dealloc(s);
    });
}

int ref_rank_fcwta_u8(const uint8_t *x, int64_t n, int64_t t, int64_t f, int64_t radius, int64_t count,
                      RefWinner *winners, int64_t *counts)
{
    return guard([&] {
        auto r = cpu_rank_fcwta(Dyn3((void *)x, Type::U8, n, t, f), radius, count, Scalar(U8(0)));
        put(r, count, winners, counts);
    });
}

int ref_rank_gather_u8(const uint8_t *in, int64_t n, int64_t t, int64_t len, uint8_t threshold, int32_t *out)
{
    return guard([&] {
        cpu_rank_gather(Dyn3((void *)in, Type::U8, n, t, len), Dyn2(out, Type::I32, n, len), Scalar(threshold));
    });
}

int ref_rank_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, int32_t *out)
{
    return guard([&] {
        cpu_rank_gather(Dyn3((void *)in, Type::F32, n, t, len), Dyn2(out, Type::I32, n, len), Scalar(threshold));
    });
}

int ref_rank_scatter_f32(const float *in, int64_t n, int64_t len, int64_t t, uint8_t *out)
{
    return guard([&] { cpu_rank_scatter(Dyn2((void *)in, Type::F32, n, len), Dyn3(out, Type::U8, n, t, len)); });
}

int ref_rank_scatter_u8(const uint8_t *in, int64_t n, int64_t len, int64_t t, uint8_t *out)
{
    return guard([&] { cpu_rank_scatter(Dyn2((void *)in, Type::U8, n, len), Dyn3(out, Type::U8, n, t, len)); });
}

int ref_rate_gather_f32(const float *in, int64_t n, int64_t t, int64_t len, float *out)
{
    return guard([&] {
        cpu_rate_gather(Dyn3((void *)in, Type::F32, n, t, len), Dyn2(out, Type::F32, n, len), Scalar(0.0f));
    });
}
}
