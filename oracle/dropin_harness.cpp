// C entry points over the reference's own public C++ API (src/spyker/opers.cpp,
// utils.cpp, impl.cpp, shape.cpp, compiled from the reference's sources where
// they lie), whose CUDA seam is this repository's HIP backend
// (spyker_amd/seam/hip_backend.cpp over libspyker_hip.so).  TEST
// INFRASTRUCTURE ONLY: it shows that the reference's own layer code runs
// unchanged on the new backend.  The reference's pybind module (bind.cpp)
// cannot be built here (it includes the CMake-generated spyker/config.h), so
// these functions stand in for its lambdas (src/bind/bind.cpp:313-406): each
// wraps caller-owned device buffers (torch tensors) as Kind::CUDA tensors with
// a no-op deleter, as bind.cpp:184-187 does, and calls the same Spyker::
// function the binding calls.
#include <spyker/opers.h>

#include <cstring>
#include <memory>
#include <string>

using namespace Spyker;

namespace
{
thread_local std::string last_error;

// A caller-owned buffer: address, device (-1 = host), spk_type / Type
// ordinal, rank and shape.
struct View
{
    int64_t ptr;
    int32_t device;
    int32_t type;
    int32_t dims;
    int32_t pad_;
    int64_t shape[6];
};

struct RefWinner
{
    int64_t c, t, z, y, x;
};

Tensor wrap(const View *v)
{
    Shape shape(v->shape, v->shape + v->dims);
    Device device = v->device < 0 ? Device(Kind::CPU) : Device(Kind::CUDA, v->device);
    return Tensor(std::shared_ptr<void>(reinterpret_cast<void *>(v->ptr), [](void *) {}), device, Type(v->type),
                  shape, false, false);
}

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
        for (size_t k = 0; k < w[s].size() && int64_t(k) < count; ++k)
            out[s * count + k] = {w[s][k].c, w[s][k].t, w[s][k].z, w[s][k].y, w[s][k].x};
    }
}
}  // namespace

extern "C" {

const char *dropin_last_error(void) { return last_error.c_str(); }

int dropin_cuda_device_count(int64_t *count)
{
    return guard([&] { *count = cudaDeviceCount(); });
}

// Conv module (bind.cpp:340-352, opers.cpp:838-889)
int dropin_conv_create(int32_t device, int64_t cin, int64_t cout, int64_t kh, int64_t kw, int64_t sy, int64_t sx,
                       int64_t pt, int64_t pl, int64_t pb, int64_t pr, void **handle)
{
    return guard([&] {
        *handle = new Conv(Device(Kind::CUDA, device), cin, cout, Expand2(kh, kw), Expand2(sy, sx),
                           Expand4(pt, pl, pb, pr));
    });
}

int dropin_conv_destroy(void *handle)
{
    return guard([&] { delete static_cast<Conv *>(handle); });
}

// the module's kernel becomes the caller's buffer (module.py sets it with wrap_array)
int dropin_conv_set_kernel(void *handle, const View *kernel)
{
    return guard([&] { static_cast<Conv *>(handle)->kernel = wrap(kernel); });
}

int dropin_conv_add_config(void *handle, double positive, double negative, int32_t stabilize, double lower,
                           double upper)
{
    return guard([&] {
        static_cast<Conv *>(handle)->stdpconfig.push_back(STDPConfig(positive, negative, stabilize != 0, lower, upper));
    });
}

int dropin_conv_forward(void *handle, const View *in, const View *out)
{
    return guard([&] { (*static_cast<Conv *>(handle))(wrap(in), wrap(out)); });
}

int dropin_conv_stdp(void *handle, const View *in, const RefWinner *win, const int64_t *counts, int64_t n,
                     int64_t maxcount, const View *out)
{
    return guard([&] {
        Winners w(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = 0; k < counts[i]; ++k)
            {
                const RefWinner &e = win[i * maxcount + k];
                w[size_t(i)].push_back(Winner{e.c, e.t, e.z, e.y, e.x});
            }
        static_cast<Conv *>(handle)->stdp(wrap(in), w, wrap(out));
    });
}

// DoG and Gabor modules on the device (bind.cpp:313-339, opers.cpp:702-772)
int dropin_dog(int32_t device, int64_t size, const double *std1, const double *std2, int64_t nfilters, int64_t pt,
               int64_t pl, int64_t pb, int64_t pr, const View *in, const View *out)
{
    return guard([&] {
        std::vector<DoGFilter> filters;
        for (int64_t i = 0; i < nfilters; ++i) filters.push_back(DoGFilter(std1[i], std2[i]));
        DoG dog(Device(Kind::CUDA, device), size, filters, Expand4(pt, pl, pb, pr));
        dog(wrap(in), wrap(out));
    });
}

int dropin_gabor(int32_t device, int64_t size, const double *params, int64_t nfilters, int64_t pt, int64_t pl,
                 int64_t pb, int64_t pr, const View *in, const View *out)
{
    return guard([&] {
        std::vector<GaborFilter> filters;
        for (int64_t i = 0; i < nfilters; ++i)
        {
            const double *p = params + 5 * i;  // sigma, theta (degrees), gamma, lambda, psi
            filters.push_back(GaborFilter(p[0], p[1], p[2], p[3], p[4]));
        }
        Gabor gabor(Device(Kind::CUDA, device), size, filters, Expand4(pt, pl, pb, pr));
        gabor(wrap(in), wrap(out));
    });
}

// free functions (bind.cpp:365-406)
int dropin_threshold(const View *x, double threshold, double value)
{
    return guard([&] { Spyker::threshold(wrap(x), Scalar(threshold), Scalar(value), true); });
}

int dropin_code(const View *in, const View *out, int64_t time, int32_t sort)
{
    return guard([&] { Spyker::code(wrap(in), wrap(out), time, sort != 0, Code::Rank); });
}

int dropin_inhibit(const View *x, double threshold)
{
    return guard([&] { Spyker::inhibit(wrap(x), Scalar(threshold), true); });
}

int dropin_convwta(const View *x, int64_t ry, int64_t rx, int64_t count, double threshold, RefWinner *win,
                   int64_t *counts)
{
    return guard([&] { put(Spyker::convwta(wrap(x), Expand2(ry, rx), count, Scalar(threshold)), count, win, counts); });
}

int dropin_fire(const View *in, const View *out, double threshold)
{
    return guard([&] { Spyker::fire(wrap(in), wrap(out), Scalar(threshold), Code::Rank); });
}

int dropin_pool(const View *in, const View *out, int64_t kh, int64_t kw, int64_t sh, int64_t sw, int64_t pt,
                int64_t pl, int64_t pb, int64_t pr)
{
    return guard([&] {
        Spyker::pool(wrap(in), wrap(out), Expand2(kh, kw), Expand2(sh, sw), Expand4(pt, pl, pb, pr));
    });
}

int dropin_pad(const View *in, const View *out, int64_t pt, int64_t pl, int64_t pb, int64_t pr, double value)
{
    return guard([&] { Spyker::pad(wrap(in), wrap(out), Expand4(pt, pl, pb, pr), Scalar(value)); });
}
}
