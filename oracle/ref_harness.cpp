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
}
