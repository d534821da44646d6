// Spyker's C++ interface for the dense STDP layer path over the C ABI
// (include/spyker_hip.h).  Shape rules, defaults and error texts follow
// src/spyker/opers.cpp and src/spyker/shape.cpp; every computation is one
// spk_* call on the default stream.  Host <-> device transfers use the HIP
// runtime directly.
#include "spyker_amd/spyker.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "spyker_hip.h"

namespace Spyker
{
namespace
{
[[noreturn]] void fail(const std::string &scope, const std::string &message)
{
    throw std::runtime_error(scope + ": " + message);
}

#define REQUIRE(cond, scope, msg)              \
    do                                         \
    {                                          \
        if (!(cond))                           \
        {                                      \
            std::ostringstream _spk_m;         \
            _spk_m << msg;                     \
            fail(scope, _spk_m.str());         \
        }                                      \
    } while (0)

void check(int32_t status, const char *scope)
{
    if (status != SPK_OK) fail(scope, spk_last_error());
}

void hip(hipError_t e, const char *scope)
{
    if (e != hipSuccess) fail(scope, hipGetErrorString(e));
}

int32_t ty(Type t) { return int32_t(t); }

// input views (opers.cpp:158-199)
Tensor atLeast(Tensor t, Size n, const char *what)
{
    Shape s = t.shape();
    if (Size(s.size()) == n - 1) s.insert(s.begin(), 1);
    REQUIRE(Size(s.size()) >= n, "Interface", "Input dimensions couldn't be viewed as at least " << what << ".");
    return t.reshape(s);
}

Tensor exactly(Tensor t, Size n, const char *what)
{
    Shape s = t.shape();
    if (Size(s.size()) == n - 1) s.insert(s.begin(), 1);
    REQUIRE(Size(s.size()) == n, "Interface", "Input dimensions couldn't be viewed as " << what << ".");
    return t.reshape(s);
}

Tensor to3(Tensor t) { return exactly(t, 3, "3D"); }
Tensor to4(Tensor t) { return exactly(t, 4, "4D"); }
Tensor to5(Tensor t) { return exactly(t, 5, "5D"); }

void checkItems(const char *name, const Shape &a, const Shape &low)
{
    for (size_t i = 0; i < a.size(); ++i) REQUIRE(a[i] >= low[i], "Interface", name << " items are incorrect.");
}

// every tensor of one call on the same GPU
void onGpu(std::initializer_list<Tensor> ts, const char *scope)
{
    const Tensor *first = nullptr;
    for (const Tensor &t : ts)
    {
        REQUIRE(t.device() == Kind::CUDA, scope,
                "This backend computes on the GPU: move the tensor with to(Device(Kind::CUDA)).");
        if (first) REQUIRE(t.device() == first->device(), scope, "Tensor devices are not compatible.");
        first = &t;
    }
}

Shape convShape(const Shape &in, const Shape &k, const Shape &stride, const Shape &pad)
{
    REQUIRE(in.size() == 5, "Shape::Conv", "Input dimensions must be 5.");
    REQUIRE(k.size() == 4, "Shape::Conv", "Kernel dimensions must be 4.");
    REQUIRE(in[2] == k[1], "Shape::Conv", "Input and kernel channels are not compatible.");
    Shape out{in[0], in[1], k[0], (in[3] + pad[0] + pad[2] - k[2]) / stride[0] + 1,
              (in[4] + pad[1] + pad[3] - k[3]) / stride[1] + 1};
    REQUIRE(out[3] > 0 && out[4] > 0, "Shape::Conv", "Input and kernel shapes are not compatible.");
    return out;
}

Shape fcShape(const Shape &in, const Shape &k)
{
    REQUIRE(in.size() == 3, "Shape::FC", "Input dimensions must be 3.");
    REQUIRE(k.size() == 2, "Shape::FC", "Kernel dimensions must be 2.");
    REQUIRE(in[2] == k[1], "Shape::FC", "Input and kernel shapes don't match.");
    return Shape{in[0], in[1], k[0]};
}

Shape filterShape(const Shape &in, Size count, Size kh, Size kw, const Shape &pad, const char *scope)
{
    Shape out{in[0], in[1] * count, in[2] + pad[0] + pad[2] - kh + 1, in[3] + pad[1] + pad[3] - kw + 1};
    REQUIRE(out[2] > 0 && out[3] > 0, scope, "Input and kernel shapes are not compatible.");
    return out;
}

Shape poolShape(const Shape &in, const Shape &k, const Shape &s, const Shape &pad)
{
    Shape out = in;
    out[3] = (in[3] + pad[0] + pad[2] - k[0]) / s[0] + 1;
    out[4] = (in[4] + pad[1] + pad[3] - k[1]) / s[1] + 1;
    REQUIRE(out[3] > 0 && out[4] > 0, "Shape::Pool", "Input and kernel shapes are not compatible.");
    return out;
}

// host winners <-> device records [n][cap] + counts [n]
struct DeviceWinners
{
    Tensor data, counts;
    Size cap;
};

DeviceWinners upload(const Winners &winners, Device device)
{
    const Size n = Size(winners.size());
    Size cap = 1;
    for (auto &w : winners) cap = std::max(cap, Size(w.size()));
    Tensor hd(Type::I64, {n, cap, 5}), hc(Type::I64, {n});
    std::memset(hd.data(), 0, size_t(hd.bytes()));
    auto *d = hd.data<int64_t>();
    auto *c = hc.data<int64_t>();
    for (Size i = 0; i < n; ++i)
    {
        c[i] = Size(winners[i].size());
        for (Size k = 0; k < c[i]; ++k)
        {
            const Winner &w = winners[i][k];
            int64_t *e = d + (i * cap + k) * 5;
            e[0] = w.c, e[1] = w.t, e[2] = w.z, e[3] = w.y, e[4] = w.x;
        }
    }
    return DeviceWinners{hd.to(device), hc.to(device), cap};
}

Winners download(const Tensor &data, const Tensor &counts)
{
    Tensor hd = data.to(Device(Kind::CPU)), hc = counts.to(Device(Kind::CPU));
    const Size n = hd.shape(0), cap = hd.shape(1);
    const auto *d = hd.data<int64_t>();
    const auto *c = hc.data<int64_t>();
    Winners out(n);
    for (Size i = 0; i < n; ++i)
        for (Size k = 0; k < c[i]; ++k)
        {
            const int64_t *e = d + (i * cap + k) * 5;
            out[i].push_back(Winner{e[0], e[1], e[2], e[3], e[4]});
        }
    return out;
}

std::vector<spk_stdp_config> configs(const std::vector<STDPConfig> &list)
{
    std::vector<spk_stdp_config> out;
    for (auto &c : list)
        out.push_back(spk_stdp_config{c.positive, c.negative, c.stabilize ? 1 : 0, c.lower, c.upper});
    return out;
}

template <typename T>
void castLoop(const void *in, Type from, T *out, Size n)
{
    switch (from)
    {
#define CASE(tp, ct)                                                          \
    case Type::tp:                                                            \
        for (Size i = 0; i < n; ++i) out[i] = T(static_cast<const ct *>(in)[i]); \
        break;
        CASE(I8, int8_t)
        CASE(I16, int16_t)
        CASE(I32, int32_t)
        CASE(I64, int64_t)
        CASE(U8, uint8_t)
        CASE(U16, uint16_t)
        CASE(U32, uint32_t)
        CASE(U64, uint64_t)
        CASE(F16, _Float16)
        CASE(F32, float)
        CASE(F64, double)
#undef CASE
    }
}

void hostCast(const void *in, Type from, void *out, Type to, Size n)
{
    switch (to)
    {
    case Type::I8: return castLoop(in, from, static_cast<int8_t *>(out), n);
    case Type::I16: return castLoop(in, from, static_cast<int16_t *>(out), n);
    case Type::I32: return castLoop(in, from, static_cast<int32_t *>(out), n);
    case Type::I64: return castLoop(in, from, static_cast<int64_t *>(out), n);
    case Type::U8: return castLoop(in, from, static_cast<uint8_t *>(out), n);
    case Type::U16: return castLoop(in, from, static_cast<uint16_t *>(out), n);
    case Type::U32: return castLoop(in, from, static_cast<uint32_t *>(out), n);
    case Type::U64: return castLoop(in, from, static_cast<uint64_t *>(out), n);
    case Type::F16: return castLoop(in, from, static_cast<_Float16 *>(out), n);
    case Type::F32: return castLoop(in, from, static_cast<float *>(out), n);
    case Type::F64: return castLoop(in, from, static_cast<double *>(out), n);
    }
}
}  // namespace

// -- types -----------------------------------------------------------------

Size typeSize(Type type)
{
    switch (type)
    {
    case Type::I8:
    case Type::U8: return 1;
    case Type::I16:
    case Type::U16:
    case Type::F16: return 2;
    case Type::I32:
    case Type::U32:
    case Type::F32: return 4;
    default: return 8;
    }
}

std::string typeName(Type type)
{
    static const char *names[] = {"i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64", "f16", "f32", "f64"};
    return names[int(type)];
}

Expand2::Expand2(const Shape &shape) : data(shape)
{
    if (shape.size() == 1) data = {shape[0], shape[0]};
    REQUIRE(data.size() == 2, "Interface::Expand2", "Incorrect input shape.");
}

Expand4::Expand4(const Shape &shape) : data(shape)
{
    if (shape.size() == 1) data = {shape[0], shape[0], shape[0], shape[0]};
    if (shape.size() == 2) data = {shape[0], shape[1], shape[0], shape[1]};
    REQUIRE(data.size() == 4, "Interface::Expand4", "Incorrect input shape.");
}

// -- tensor ------------------------------------------------------------------

Tensor::Tensor() {}

Tensor::Tensor(Shared data, Size offset, Device device, Type type, Shape shape)
    : _data(data), _offset(offset), _device(device), _type(type), _shape(shape)
{
}

Tensor::Tensor(Type type, const Shape &shape) : Tensor(Device(Kind::CPU), type, shape) {}

Tensor::Tensor(Device device, Type type, const Shape &shape) : _device(device), _type(type), _shape(shape)
{
    for (Size d : shape) REQUIRE(d >= 0, "Interface::Tensor", "Tensor dimensions can't be negative.");
    const size_t bytes = size_t(std::max<Size>(numel() * typeSize(type), 1));
    if (device == Kind::CPU)
    {
        void *p = std::malloc(bytes);
        REQUIRE(p, "Interface::Tensor", "Host allocation of " << bytes << " bytes failed.");
        _data = Shared(p, std::free);
    }
    else
    {
        void *p = nullptr;
        check(spk_alloc(&p, int64_t(bytes)), "Interface::Tensor");
        _data = Shared(p, [](void *q) { spk_dealloc(q); });
    }
}

Tensor::Tensor(Shared data, Device device, Type type, const Shape &shape)
    : _data(data), _device(device), _type(type), _shape(shape)
{
}

void *Tensor::data() const { return static_cast<char *>(_data.get()) + _offset; }

Size Tensor::numel() const
{
    Size n = 1;
    for (Size d : _shape) n *= d;
    return n;
}

Size Tensor::shape(Size index) const
{
    if (index < 0) index += dims();
    REQUIRE(index >= 0 && index < dims(), "Interface::Tensor", "Index is out of range.");
    return _shape[size_t(index)];
}

Tensor Tensor::copy() const
{
    Tensor out(_device, _type, _shape);
    out.from(*this);
    return out;
}

Tensor Tensor::to(Type type) const
{
    if (type == _type) return *this;
    Tensor out(_device, type, _shape);
    if (_device == Kind::CPU)
        hostCast(data(), _type, out.data(), type, numel());
    else
        check(spk_cast(data(), ty(_type), out.data(), ty(type), numel(), nullptr), "Interface::Tensor");
    return out;
}

Tensor Tensor::to(Device device) const
{
    if (device == _device && device.kind() == _device.kind()) return *this;
    Tensor out(device, _type, _shape);
    to(out);
    return out;
}

void Tensor::to(Tensor other) const
{
    REQUIRE(other.bytes() == bytes(), "Interface::Tensor", "Tensor sizes are not compatible.");
    if (bytes() == 0) return;
    if (_device == Kind::CPU && other.device() == Kind::CPU)
        std::memcpy(other.data(), data(), size_t(bytes()));
    else
        hip(hipMemcpy(other.data(), data(), size_t(bytes()), hipMemcpyDefault), "Interface::Tensor");
}

void Tensor::from(Tensor tensor) const { tensor.to(*this); }

Tensor Tensor::reshape(Shape shape) const
{
    Size known = 1, free = -1;
    for (size_t i = 0; i < shape.size(); ++i)
    {
        if (shape[i] == -1)
        {
            REQUIRE(free < 0, "Interface::Tensor", "Only one dimension can be inferred.");
            free = Size(i);
        }
        else
            known *= shape[i];
    }
    if (free >= 0) shape[size_t(free)] = known ? numel() / known : 0;
    Size total = 1;
    for (Size d : shape) total *= d;
    REQUIRE(total == numel(), "Interface::Tensor", "Tensor can't be reshaped to the given shape.");
    return Tensor(_data, _offset, _device, _type, shape);
}

Tensor &Tensor::fill(F64 value)
{
    Tensor host(Type::F64, {numel()});
    for (Size i = 0; i < numel(); ++i) host.data<double>()[i] = value;
    Tensor typed = host.to(_type);
    typed.to(*this);
    return *this;
}

Tensor Tensor::operator[](Size index) const
{
    REQUIRE(dims() > 0 && index >= 0 && index < _shape[0], "Interface::Tensor", "Index is out of range.");
    Shape sub(_shape.begin() + 1, _shape.end());
    Size inner = 1;
    for (Size d : sub) inner *= d;
    return Tensor(_data, _offset + index * inner * typeSize(_type), _device, _type, sub);
}

// -- control ----------------------------------------------------------------

void randomSeed(Size seed) { check(spk_random_seed(seed), "Interface::RandomSeed"); }

bool cudaAvailable()
{
    int32_t n = 0;
    return spk_device_count(&n) == SPK_OK && n > 0;
}

Size cudaDeviceCount()
{
    int32_t n = 0;
    check(spk_device_count(&n), "Interface::Device");
    return n;
}

void cudaSetDevice(Size index) { check(spk_set_device(int32_t(index)), "Interface::Device"); }

Size cudaMemoryTaken(Size)
{
    int64_t b = 0;
    check(spk_memory_taken(&b), "Interface::Memory");
    return b;
}

Size cudaMemoryUsed(Size)
{
    int64_t b = 0;
    check(spk_memory_used(&b), "Interface::Memory");
    return b;
}

void cudaCacheClear(Size) { check(spk_cache_clear(), "Interface::Memory"); }

// -- functions (opers.cpp) ---------------------------------------------------

Tensor dog(Tensor input, Tensor kernel, Expand4 pad)
{
    input = to4(input);
    REQUIRE(kernel.dims() == 4 && kernel.shape(0) == 2, "Shape::DoG", "Kernel shape is invalid.");
    Tensor output(input.device(), kernel.type(),
                  filterShape(input.shape(), kernel.shape(1), kernel.shape(2), kernel.shape(3), pad.get(),
                              "Shape::DoG"));
    return dog(input, kernel, output, pad);
}

Tensor dog(Tensor input, Tensor kernel, Tensor output, Expand4 pad)
{
    input = to4(input).to(kernel.type()), output = to4(output);
    onGpu({input, kernel, output}, "Interface::DoG");
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    REQUIRE(kernel.type() == output.type(), "Interface::DoG", "Kernel and output must have the same type.");
    REQUIRE(kernel.dims() == 4 && kernel.shape(0) == 2, "Shape::DoG", "Kernel shape is invalid.");
    REQUIRE(output.shape() == filterShape(input.shape(), kernel.shape(1), kernel.shape(2), kernel.shape(3), pad.get(),
                                          "Shape::DoG"),
            "Interface::DoG", "Tensor shapes are not compatible.");
    REQUIRE(kernel.shape(2) == kernel.shape(3), "Interface::DoG", "Kernel must be square.");
    check(spk_dog(input.data(), kernel.data(), output.data(), input.shape(0) * input.shape(1), input.shape(2),
                  input.shape(3), kernel.shape(1), kernel.shape(2), pad[0], pad[1], pad[2], pad[3], nullptr),
          "Interface::DoG");
    return output;
}

Tensor gabor(Tensor input, Tensor kernel, Expand4 pad)
{
    input = to4(input);
    REQUIRE(kernel.dims() == 3, "Shape::Gabor", "Kernel dimensions must be 3.");
    Tensor output(input.device(), kernel.type(),
                  filterShape(input.shape(), kernel.shape(0), kernel.shape(1), kernel.shape(2), pad.get(),
                              "Shape::Gabor"));
    return gabor(input, kernel, output, pad);
}

Tensor gabor(Tensor input, Tensor kernel, Tensor output, Expand4 pad)
{
    input = to4(input).to(kernel.type()), output = to4(output);
    onGpu({input, kernel, output}, "Interface::Gabor");
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    REQUIRE(kernel.type() == output.type(), "Interface::Gabor", "Kernel and output must have the same type.");
    REQUIRE(kernel.dims() == 3, "Shape::Gabor", "Kernel dimensions must be 3.");
    REQUIRE(output.shape() == filterShape(input.shape(), kernel.shape(0), kernel.shape(1), kernel.shape(2),
                                          pad.get(), "Shape::Gabor"),
            "Interface::Gabor", "Tensor shapes are not compatible.");
    REQUIRE(kernel.shape(1) == kernel.shape(2), "Interface::Gabor", "Kernel must be square.");
    check(spk_gabor(input.data(), kernel.data(), output.data(), input.shape(0) * input.shape(1), input.shape(2),
                    input.shape(3), kernel.shape(0), kernel.shape(1), pad[0], pad[1], pad[2], pad[3], nullptr),
          "Interface::Gabor");
    return output;
}

Tensor conv(Tensor input, Tensor kernel, Expand2 stride, Expand4 pad)
{
    input = to5(input);
    checkItems("Stride", stride.get(), {1, 1}), checkItems("Pad", pad.get(), {0, 0, 0, 0});
    Tensor output(input.device(), kernel.type(), convShape(input.shape(), kernel.shape(), stride.get(), pad.get()));
    return conv(input, kernel, output, stride, pad);
}

Tensor conv(Tensor input, Tensor kernel, Tensor output, Expand2 stride, Expand4 pad)
{
    // U8 spikes go to the kernel as they are: the conversion to the kernel's
    // type (opers.cpp:410) is exact and happens inside the kernel's loads
    input = to5(input), output = to5(output);
    if (input.type() != Type::U8) input = input.to(kernel.type());
    onGpu({input, kernel, output}, "Interface::Conv");
    checkItems("Stride", stride.get(), {1, 1}), checkItems("Pad", pad.get(), {0, 0, 0, 0});
    REQUIRE(kernel.type() == output.type(), "Interface::Conv", "Kernel and output must have the same type.");
    REQUIRE(output.shape() == convShape(input.shape(), kernel.shape(), stride.get(), pad.get()), "Interface::Conv",
            "Tensor shapes are not compatible.");
    check(spk_conv(input.data(), ty(input.type()), kernel.data(), ty(kernel.type()), output.data(), ty(output.type()),
                   input.shape(0) * input.shape(1), input.shape(2), input.shape(3), input.shape(4), kernel.shape(0),
                   kernel.shape(2), kernel.shape(3), stride[0], stride[1], pad[0], pad[1], pad[2], pad[3], nullptr),
          "Interface::Conv");
    return output;
}

Tensor fc(Tensor input, Tensor kernel)
{
    input = to3(input);
    Tensor output(input.device(), kernel.type(), fcShape(input.shape(), kernel.shape()));
    return fc(input, kernel, output);
}

Tensor fc(Tensor input, Tensor kernel, Tensor output)
{
    input = to3(input), output = to3(output);
    if (input.type() != Type::U8) input = input.to(kernel.type());
    onGpu({input, kernel, output}, "Interface::FC");
    REQUIRE(kernel.type() == output.type(), "Interface::FC", "Kernel and output must have the same type.");
    REQUIRE(output.shape() == fcShape(input.shape(), kernel.shape()), "Interface::FC",
            "Tensor shapes are not compatible.");
    check(spk_fc(input.data(), ty(input.type()), kernel.data(), output.data(), input.shape(0) * input.shape(1),
                 input.shape(2), kernel.shape(0), nullptr),
          "Interface::FC");
    return output;
}

Tensor pad(Tensor input, Expand4 pad, F64 value)
{
    input = to5(input);
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    Shape s = input.shape();
    s[3] += pad[0] + pad[2], s[4] += pad[1] + pad[3];
    Tensor output(input.device(), input.type(), s);
    return Spyker::pad(input, output, pad, value);
}

Tensor pad(Tensor input, Tensor output, Expand4 pad, F64 value)
{
    input = to5(input), output = to5(output);
    onGpu({input, output}, "Interface::Pad");
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    REQUIRE(input.type() == output.type(), "Interface::Pad", "Input and output must have the same data type.");
    Shape s = input.shape();
    s[3] += pad[0] + pad[2], s[4] += pad[1] + pad[3];
    REQUIRE(output.shape() == s, "Interface::Pad", "Tensor shapes are not compatible.");
    check(spk_pad(input.data(), output.data(), ty(input.type()), input.shape(0) * input.shape(1) * input.shape(2),
                  input.shape(3), input.shape(4), pad[0], pad[1], pad[2], pad[3], value, nullptr),
          "Interface::Pad");
    return output;
}

Tensor threshold(Tensor input, F64 threshold, F64 value, bool inplace)
{
    onGpu({input}, "Interface::Threshold");
    if (!inplace) input = input.copy();
    check(spk_threshold(input.data(), ty(input.type()), input.numel(), threshold, value, nullptr),
          "Interface::Threshold");
    return input;
}

Tensor code(Tensor input, Size time, bool sort, Type type, Code code)
{
    input = atLeast(input, 2, "2D");
    REQUIRE(time > 0, "Interface::RankCode", "Input time can't be less than 1.");
    Shape s = input.shape();
    s.insert(s.begin() + 1, time);
    Tensor output(input.device(), type, s);
    return Spyker::code(input, output, time, sort, code);
}

Tensor code(Tensor input, Tensor output, Size time, bool sort, Code code)
{
    input = atLeast(input, 2, "2D"), output = atLeast(output, 3, "3D");
    onGpu({input, output}, "Interface::RankCode");
    REQUIRE(time > 0, "Interface::RankCode", "Input time can't be less than 1.");
    Shape s = input.shape();
    s.insert(s.begin() + 1, time);
    REQUIRE(output.shape() == s, "Interface::RankCode", "Tensor shapes are not compatible.");
    const Size n = input.shape(0);
    auto coder = code == Code::Rank ? spk_rank_code : spk_rate_code;
    check(coder(input.data(), ty(input.type()), output.data(), ty(output.type()), n, input.numel() / n, time,
                sort ? 1 : 0, nullptr),
          "Interface::RankCode");
    return output;
}

Tensor fire(Tensor input, F64 threshold, Type type, Code code)
{
    Tensor output(input.device(), type, input.shape());
    return fire(input, output, threshold, code);
}

Tensor fire(Tensor input, Tensor output, F64 threshold, Code code)
{
    onGpu({input, output}, "Interface::Fire");
    REQUIRE(output.shape() == input.shape(), "Interface::Fire", "Tensor shapes are not compatible.");
    if (code == Code::Rank)
    {
        check(spk_rank_fire(input.data(), ty(input.type()), output.data(), ty(output.type()), input.numel(),
                            threshold, nullptr),
              "Interface::Fire");
        return output;
    }
    // opers.cpp:540-550: [N][T][...] as [N][T][-1]
    REQUIRE(input.dims() >= 3, "Interface::Fire", "Input dimensions couldn't be viewed as at least 3D.");
    check(spk_rate_fire(input.data(), ty(input.type()), output.data(), ty(output.type()), input.shape(0),
                        input.shape(1), input.numel() / (input.shape(0) * input.shape(1)), threshold, nullptr),
          "Interface::Fire");
    return output;
}

Tensor pool(Tensor input, Expand2 kernel, Expand2 stride, Expand4 pad)
{
    input = to5(input);
    if (stride.get() == Shape{0, 0}) stride = kernel;
    checkItems("Kernel", kernel.get(), {1, 1}), checkItems("Stride", stride.get(), {1, 1});
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    Tensor output(input.device(), input.type(), poolShape(input.shape(), kernel.get(), stride.get(), pad.get()));
    return pool(input, output, kernel, stride, pad);
}

Tensor pool(Tensor input, Tensor output, Expand2 kernel, Expand2 stride, Expand4 pad)
{
    input = to5(input), output = to5(output);
    if (stride.get() == Shape{0, 0}) stride = kernel;
    onGpu({input, output}, "Interface::Pool");
    checkItems("Kernel", kernel.get(), {1, 1}), checkItems("Stride", stride.get(), {1, 1});
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    REQUIRE(input.type() == output.type(), "Interface::Pool", "Input and output must have the same data type.");
    REQUIRE(output.shape() == poolShape(input.shape(), kernel.get(), stride.get(), pad.get()), "Interface::Pool",
            "Tensor shapes are not compatible.");
    check(spk_rank_pool(input.data(), output.data(), ty(input.type()),
                        input.shape(0) * input.shape(1) * input.shape(2), input.shape(3), input.shape(4), kernel[0],
                        kernel[1], stride[0], stride[1], pad[0], pad[1], pad[2], pad[3], nullptr),
          "Interface::Pool");
    return output;
}

Tensor inhibit(Tensor input, F64 threshold, bool inplace)
{
    onGpu({input}, "Interface::Inhibit");
    if (!inplace) input = input.copy();
    Tensor in = to5(input);
    check(spk_rank_inhibit(in.data(), ty(in.type()), in.shape(0), in.shape(1), in.shape(2),
                           in.shape(3) * in.shape(4), threshold, nullptr),
          "Interface::Inhibit");
    return input;
}

Winners fcwta(Tensor input, Size radius, Size count, F64 threshold)
{
    input = to3(input);
    onGpu({input}, "Interface::FCWTA");
    REQUIRE(radius >= 0, "Interface::FCWTA", "Radius can't be less than 0.");
    REQUIRE(count > 0, "Interface::FCWTA", "Winners count can't be less than 1.");
    const Size n = input.shape(0);
    Tensor data(input.device(), Type::I64, {n, count, 5}), counts(input.device(), Type::I64, {n});
    check(spk_rank_fcwta(input.data(), ty(input.type()), n, input.shape(1), input.shape(2), radius, count, threshold,
                         data.data<spk_winner>(), counts.data<int64_t>(), nullptr),
          "Interface::FCWTA");
    return download(data, counts);
}

Winners convwta(Tensor input, Expand2 radius, Size count, F64 threshold)
{
    input = to5(input);
    onGpu({input}, "Interface::ConvWTA");
    checkItems("Radius", radius.get(), {0, 0});
    REQUIRE(count > 0, "Interface::ConvWTA", "Winners count can't be less than 1.");
    const Size n = input.shape(0);
    Tensor data(input.device(), Type::I64, {n, count, 5}), counts(input.device(), Type::I64, {n});
    check(spk_rank_convwta(input.data(), ty(input.type()), n, input.shape(1), input.shape(2), input.shape(3),
                           input.shape(4), radius[0], radius[1], count, threshold, data.data<spk_winner>(),
                           counts.data<int64_t>(), nullptr),
          "Interface::ConvWTA");
    return download(data, counts);
}

// -- modules (opers.cpp:702-942) ---------------------------------------------

DoG::DoG(Size size, const std::vector<DoGFilter> &filters, Expand4 pad, Type type)
    : DoG(Device(Kind::CPU), size, filters, pad, type)
{
}

DoG::DoG(Device device, Size size, const std::vector<DoGFilter> &filters, Expand4 pad, Type type)
    : _init(true), _device(device), _pad(pad)
{
    REQUIRE(size > 0, "Interface::DoG", "Input size can't be less than 1.");
    REQUIRE(!filters.empty(), "Interface::DoG", "Input filters count can't be less then 1.");
    checkItems("Pad", _pad.get(), {0, 0, 0, 0});
    const Size count = Size(filters.size()), k = 2 * size + 1;
    Tensor host(Type::F64, {2, count, k, k});
    for (Size i = 0; i < count; ++i)
    {
        check(spk_gaussian_kernel(host[0][i].data<double>(), k, filters[size_t(i)].std1), "Interface::DoG");
        check(spk_gaussian_kernel(host[1][i].data<double>(), k, filters[size_t(i)].std2), "Interface::DoG");
    }
    kernel = host.to(type).to(_device);
}

Tensor DoG::operator()(Tensor input, Tensor output)
{
    REQUIRE(_init, "Interface::DoG", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::DoG", "Input device is not the same as layer's.");
    return dog(input, kernel, output, _pad);
}

Tensor DoG::operator()(Tensor input)
{
    REQUIRE(_init, "Interface::DoG", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::DoG", "Input device is not the same as layer's.");
    return dog(input, kernel, _pad);
}

LoG::LoG(Size size, const std::vector<F64> &stds, Expand4 pad, Type type)
    : LoG(Device(Kind::CPU), size, stds, pad, type)
{
}

LoG::LoG(Device device, Size size, const std::vector<F64> &stds, Expand4 pad, Type type)
    : _init(true), _device(device), _pad(pad)
{
    REQUIRE(size > 0, "Interface::LoG", "Input size can't be less than 1.");
    REQUIRE(!stds.empty(), "Interface::LoG", "Input stds count can't be less then 1.");
    checkItems("Pad", _pad.get(), {0, 0, 0, 0});
    const Size count = Size(stds.size()), k = 2 * size + 1;
    Tensor host(Type::F64, {2, count, k, k});
    for (Size i = 0; i < count; ++i)
    {
        check(spk_gaussian_kernel(host[0][i].data<double>(), k, stds[size_t(i)] / std::sqrt(2)), "Interface::LoG");
        check(spk_gaussian_kernel(host[1][i].data<double>(), k, stds[size_t(i)] * std::sqrt(2)), "Interface::LoG");
    }
    kernel = host.to(type).to(_device);
}

Tensor LoG::operator()(Tensor input, Tensor output)
{
    REQUIRE(_init, "Interface::LoG", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::LoG", "Input device is not the same as layer's.");
    input = to4(input).to(kernel.type()), output = to4(output);
    onGpu({input, kernel, output}, "Interface::LoG");
    REQUIRE(kernel.type() == Type::F32 && output.type() == Type::F32, "Interface::LoG",
            "Kernel and output must be F32 on this backend.");
    Shape s = filterShape(input.shape(), 2 * kernel.shape(1), kernel.shape(2), kernel.shape(3), _pad.get(),
                          "Shape::LoG");
    REQUIRE(output.shape() == s, "Interface::LoG", "Tensor shapes are not compatible.");
    check(spk_log(input.data(), kernel.data(), output.data(), input.shape(0) * input.shape(1), input.shape(2),
                  input.shape(3), kernel.shape(1), kernel.shape(2), _pad[0], _pad[1], _pad[2], _pad[3], nullptr),
          "Interface::LoG");
    return output;
}

Tensor LoG::operator()(Tensor input)
{
    REQUIRE(_init, "Interface::LoG", "Layer is used before initializing.");
    Tensor in = to4(input);
    Tensor output(in.device(), kernel.type(),
                  filterShape(in.shape(), 2 * kernel.shape(1), kernel.shape(2), kernel.shape(3), _pad.get(),
                              "Shape::LoG"));
    return (*this)(in, output);
}

ZCA &ZCA::fit(Tensor input, F64 epsilon, bool transform_)
{
    REQUIRE(input.dims() >= 2, "Interface::ZCA", "Input dimensions must be at least 2D.");
    onGpu({input}, "Interface::ZCA");
    REQUIRE(input.type() == Type::F32, "Interface::ZCA", "Input data type must be F32.");
    const Size rows = input.shape(0), cols = input.numel() / rows;
    mean = Tensor(input.device(), Type::F32, {cols});
    transform = Tensor(input.device(), Type::F32, {cols, cols});
    check(spk_zca_fit(input.data(), mean.data(), transform.data(), rows, cols, epsilon, transform_ ? 1 : 0, nullptr),
          "Interface::ZCA");
    return *this;
}

Tensor ZCA::operator()(Tensor input, bool inplace)
{
    REQUIRE(bool(mean), "Interface::ZCA", "Layer is not initialized.");
    REQUIRE(bool(transform), "Interface::ZCA", "Layer is not initialized.");
    if (!inplace) input = input.copy();
    Tensor in = atLeast(input, 2, "2D");
    onGpu({in, mean, transform}, "Interface::ZCA");
    REQUIRE(in.type() == Type::F32, "Interface::ZCA", "Input data type must be F32.");
    const Size rows = in.shape(0), cols = rows ? in.numel() / rows : 0;
    REQUIRE(mean.numel() == cols && transform.shape() == Shape({cols, cols}), "Shape::ZCA",
            "Input and mean shapes are not compatible.");
    check(spk_zca_trans(in.data(), mean.data(), transform.data(), rows, cols, nullptr), "Interface::ZCA");
    return input;
}

Tensor ZCA::split(Tensor input, Tensor output)
{
    Tensor in = atLeast(input, 2, "2D");
    onGpu({in, output}, "Interface::ZCA");
    REQUIRE(in.type() == output.type(), "Interface::ZCA", "Input and output type must have the same type.");
    Shape s = in.shape();
    s.insert(s.begin() + 1, 2);
    REQUIRE(output.numel() == in.numel() * 2 && output.shape(0) == in.shape(0), "Interface::ZCA",
            "Output shape is incorrect.");
    const Size rows = in.shape(0), cols = rows ? in.numel() / rows : 0;
    check(spk_zca_split(in.data(), output.data(), ty(in.type()), rows, cols, nullptr), "Interface::ZCA");
    return output;
}

Tensor ZCA::split(Tensor input)
{
    Tensor in = atLeast(input, 2, "2D");
    Shape s = in.shape();
    s.insert(s.begin() + 1, 2);
    return split(in, Tensor(in.device(), in.type(), s));
}

Gabor::Gabor(Size size, const std::vector<GaborFilter> &filters, Expand4 pad, Type type)
    : Gabor(Device(Kind::CPU), size, filters, pad, type)
{
}

Gabor::Gabor(Device device, Size size, const std::vector<GaborFilter> &filters, Expand4 pad, Type type)
    : _init(true), _device(device), _pad(pad)
{
    REQUIRE(size > 0, "Interface::Gabor", "Input size can't be less than 1.");
    REQUIRE(!filters.empty(), "Interface::Gabor", "Input filters count can't be less then 1.");
    checkItems("Pad", _pad.get(), {0, 0, 0, 0});
    const Size count = Size(filters.size()), k = 2 * size + 1;
    Tensor host(Type::F64, {count, k, k});
    for (Size i = 0; i < count; ++i)
    {
        const GaborFilter &f = filters[size_t(i)];
        check(spk_gabor_kernel(host[i].data<double>(), k, f.sigma, f.theta, f.gamma, f.lambda, f.psi),
              "Interface::Gabor");
    }
    kernel = host.to(type).to(_device);
}

Tensor Gabor::operator()(Tensor input, Tensor output)
{
    REQUIRE(_init, "Interface::Gabor", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Gabor", "Input device is not the same as layer's.");
    return gabor(input, kernel, output, _pad);
}

Tensor Gabor::operator()(Tensor input)
{
    REQUIRE(_init, "Interface::Gabor", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Gabor", "Input device is not the same as layer's.");
    return gabor(input, kernel, _pad);
}

Conv::Conv(Size input, Size output, Expand2 kernel, Expand2 stride, Expand4 pad, F64 mean, F64 std, Type type)
    : Conv(Device(Kind::CPU), input, output, kernel, stride, pad, mean, std, type)
{
}

Conv::Conv(Device device, Size input, Size output, Expand2 k, Expand2 stride, Expand4 pad, F64 mean, F64 std,
           Type type)
    : _init(true), _device(device), _stride(stride), _pad(pad)
{
    REQUIRE(input > 0, "Interface::Conv", "Input size can't be less than 1.");
    REQUIRE(output > 0, "Interface::Conv", "Output size can't be less than 1.");
    checkItems("Kernel", k.get(), {1, 1}), checkItems("Stride", _stride.get(), {1, 1});
    checkItems("Pad", _pad.get(), {0, 0, 0, 0});
    Tensor host(Type::F64, {output, input, k[0], k[1]});
    check(spk_normal_kernel(host.data<double>(), host.numel(), mean, std), "Interface::Conv");
    kernel = host.to(type).to(_device);
}

Tensor Conv::operator()(Tensor input, Tensor output)
{
    REQUIRE(_init, "Interface::Conv", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Conv", "Input device is not the same as layer's.");
    return conv(input, kernel, output, _stride, _pad);
}

Tensor Conv::operator()(Tensor input)
{
    REQUIRE(_init, "Interface::Conv", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Conv", "Input device is not the same as layer's.");
    return conv(input, kernel, _stride, _pad);
}

void Conv::stdp(Tensor input, const Winners &winners, Tensor output)
{
    REQUIRE(_init, "Interface::STDP", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::STDP", "Input device is not the same as layer's.");
    REQUIRE(!stdpconfig.empty(), "Interface::STDP", "At least one STDP Config must be given.");
    REQUIRE(!winners.empty(), "Interface::STDP", "Winners can't be empty.");
    input = to5(input), output = to5(output);
    onGpu({input, kernel, output}, "Interface::STDP");
    REQUIRE(input.type() == output.type(), "Interface::STDP", "Input and output must have the same data type.");
    REQUIRE(output.shape() == convShape(input.shape(), kernel.shape(), _stride.get(), _pad.get()), "Interface::STDP",
            "Tensor shapes are not compatible.");
    REQUIRE(Size(winners.size()) == input.shape(0), "Interface::STDP", "Winners and input batch sizes differ.");
    const DeviceWinners w = upload(winners, input.device());
    const auto cfg = configs(stdpconfig);
    check(spk_rank_convstdp(input.data(), kernel.data(), ty(kernel.type()), output.data(), input.shape(0),
                            input.shape(1), input.shape(2), input.shape(3), input.shape(4), kernel.shape(0),
                            kernel.shape(2), kernel.shape(3), output.shape(3), output.shape(4), cfg.data(),
                            Size(cfg.size()), w.data.data<spk_winner>(), w.counts.data<int64_t>(), w.cap, _stride[0],
                            _stride[1], _pad[0], _pad[1], _pad[2], _pad[3], nullptr),
          "Interface::STDP");
}

FC::FC(Size input, Size output, F64 mean, F64 std, Type type) : FC(Device(Kind::CPU), input, output, mean, std, type)
{
}

FC::FC(Device device, Size input, Size output, F64 mean, F64 std, Type type) : _init(true), _device(device)
{
    REQUIRE(input > 0, "Interface::FC", "Input size can't be less than 1.");
    REQUIRE(output > 0, "Interface::FC", "Output size can't be less than 1.");
    Tensor host(Type::F64, {output, input});
    check(spk_normal_kernel(host.data<double>(), host.numel(), mean, std), "Interface::FC");
    kernel = host.to(type).to(_device);
}

Tensor FC::operator()(Tensor input, Tensor output)
{
    REQUIRE(_init, "Interface::FC", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::FC", "Input device is not the same as layer's.");
    return fc(input, kernel, output);
}

Tensor FC::operator()(Tensor input)
{
    REQUIRE(_init, "Interface::FC", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::FC", "Input device is not the same as layer's.");
    return fc(input, kernel);
}

void FC::stdp(Tensor input, const Winners &winners, Tensor output)
{
    input = to3(input), output = to3(output);
    REQUIRE(!stdpconfig.empty(), "Interface::STDP", "At least one STDP Config must be given.");
    REQUIRE(!winners.empty(), "Interface::STDP", "Winners can't be empty.");
    REQUIRE(_init, "Interface::STDP", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::STDP", "Input device is not the same as layer's.");
    onGpu({input, kernel, output}, "Interface::STDP");
    REQUIRE(input.type() == output.type(), "Interface::STDP", "Input and output must have the same data type.");
    REQUIRE(output.shape() == fcShape(input.shape(), kernel.shape()), "Interface::STDP",
            "Tensor shapes are not compatible.");
    REQUIRE(Size(winners.size()) == input.shape(0), "Interface::STDP", "Winners and input batch sizes differ.");
    const DeviceWinners w = upload(winners, input.device());
    const auto cfg = configs(stdpconfig);
    check(spk_rank_fcstdp(input.data(), kernel.data(), ty(kernel.type()), output.data(), input.shape(0),
                          input.shape(1), input.shape(2), kernel.shape(0), cfg.data(), Size(cfg.size()),
                          w.data.data<spk_winner>(), w.counts.data<int64_t>(), w.cap, nullptr),
          "Interface::STDP");
}

// -- readout (opers.cpp:554-598) ---------------------------------------------

Tensor gather(Tensor input, F64 threshold, Type type, Code code)
{
    input = atLeast(input, 3, "3D");
    Shape s = input.shape();
    s.erase(s.begin() + 1);
    return gather(input, Tensor(input.device(), type, s), threshold, code);
}

Tensor gather(Tensor input, Tensor output, F64 threshold, Code code)
{
    input = atLeast(input, 3, "3D"), output = atLeast(output, 2, "2D");
    onGpu({input, output}, "Interface::Gather");
    Shape s = input.shape();
    s.erase(s.begin() + 1);
    REQUIRE(output.shape() == s, "Interface::Gather", "Tensor shapes are not compatible.");
    const Size n = input.shape(0), t = input.shape(1), len = n ? output.numel() / n : 0;
    if (code == Code::Rank)
        check(spk_rank_gather(input.data(), ty(input.type()), output.data(), ty(output.type()), n, t, len, threshold,
                              nullptr),
              "Interface::Gather");
    else
        check(spk_rate_gather(input.data(), ty(input.type()), output.data(), ty(output.type()), n, t, len, nullptr),
              "Interface::Gather");
    return output;
}

Tensor scatter(Tensor input, Size time, Type type)
{
    input = atLeast(input, 2, "2D");
    REQUIRE(time > 0, "Interface::Scatter", "Input time can't be less than 1.");
    Shape s = input.shape();
    s.insert(s.begin() + 1, time);
    return scatter(input, Tensor(input.device(), type, s));
}

Tensor scatter(Tensor input, Tensor output)
{
    input = atLeast(input, 2, "2D"), output = atLeast(output, 3, "3D");
    onGpu({input, output}, "Interface::Scatter");
    Shape s = input.shape();
    s.insert(s.begin() + 1, output.shape(1));
    REQUIRE(output.shape() == s, "Interface::Scatter", "Tensor shapes are not compatible.");
    const Size n = input.shape(0), len = n ? input.numel() / n : 0;
    check(spk_rank_scatter(input.data(), ty(input.type()), output.data(), ty(output.type()), n, len, output.shape(1),
                           nullptr),
          "Interface::Scatter");
    return output;
}

// -- sparse event path (utils.cpp:206-257, opers.cpp:866-889,982-1043) ----------

namespace
{
std::shared_ptr<void> newSparse()
{
    return std::shared_ptr<void>(new spk_sparse{}, [](void *p) {
        auto *d = static_cast<spk_sparse *>(p);
        (void)spk_sparse_release(d);
        delete d;
    });
}

const spk_sparse &desc(const SparseTensor &t, const char *scope)
{
    REQUIRE(bool(t), scope, "Sparse can't be used without initialization.");
    return *static_cast<const spk_sparse *>(t.data());
}

spk_sparse *out(std::shared_ptr<void> &d) { return static_cast<spk_sparse *>(d.get()); }
}  // namespace

SparseTensor::SparseTensor() {}

SparseTensor::SparseTensor(std::shared_ptr<void> descriptor, Device device) : _data(descriptor), _device(device) {}

SparseTensor::SparseTensor(Tensor tensor, F64 threshold) : _device(tensor.device())
{
    tensor = exactly(tensor, 5, "5D");
    onGpu({tensor}, "Interface::Sparse");
    auto d = newSparse();
    check(spk_sparse_from_dense(tensor.data(), ty(tensor.type()), tensor.shape(0), tensor.shape(1), tensor.shape(2),
                                tensor.shape(3), tensor.shape(4), threshold, out(d), nullptr),
          "Interface::Sparse");
    _data = d;
}

Size SparseTensor::dims() const { return _data ? 5 : 0; }

Size SparseTensor::numel() const { return desc(*this, "Interface::Sparse").total; }

Size SparseTensor::bytes() const
{
    const spk_sparse &d = desc(*this, "Interface::Sparse");
    return (d.n * d.t + 1) * 8 + d.total * 4;
}

Shape SparseTensor::shape() const
{
    if (!_data) return {};
    const spk_sparse &d = desc(*this, "Interface::Sparse");
    return {d.n, d.t, d.c, d.h, d.w};
}

Size SparseTensor::shape(Size index) const
{
    const Shape s = shape();
    REQUIRE(index >= 0 && index < Size(s.size()), "Interface::Sparse", "Given index is out of range.");
    return s[size_t(index)];
}

F64 SparseTensor::sparsity() const
{
    const Shape s = shape();
    Size total = 1;
    for (Size v : s) total *= v;
    return F64(numel()) / F64(total);
}

SparseTensor SparseTensor::copy() const { return Sparse::pad(*this, 0); }

Tensor SparseTensor::dense() const
{
    const spk_sparse &d = desc(*this, "Interface::Sparse");
    Tensor output(_device, Type::U8, shape());
    check(spk_sparse_to_dense(&d, output.data(), nullptr), "Interface::Sparse");
    return output;
}

namespace Sparse
{
SparseTensor conv(SparseTensor input, Tensor kernel, F64 threshold, Expand2 stride, Expand4 pad)
{
    const spk_sparse &d = desc(input, "Interface::Conv");
    checkItems("Stride", stride.get(), {1, 1}), checkItems("Pad", pad.get(), {0, 0, 0, 0});
    kernel = to4(kernel);
    onGpu({kernel}, "Interface::Conv");
    REQUIRE(kernel.shape(1) == d.c, "Interface::Conv", "Input and kernel shapes are not compatible.");
    auto o = newSparse();
    check(spk_sparse_conv(&d, kernel.data(), ty(kernel.type()), kernel.shape(0), kernel.shape(2), kernel.shape(3),
                          threshold, stride[0], stride[1], pad[0], pad[1], pad[2], pad[3], out(o), nullptr),
          "Interface::Conv");
    return SparseTensor(o, input.device());
}

SparseTensor pad(SparseTensor input, Expand4 pad)
{
    const spk_sparse &d = desc(input, "Interface::Pad");
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    auto o = newSparse();
    check(spk_sparse_pad(&d, pad[0], pad[1], pad[2], pad[3], out(o), nullptr), "Interface::Pad");
    return SparseTensor(o, input.device());
}

SparseTensor code(Tensor input, Size time, bool sort)
{
    if (input.dims() == 3) input = input.reshape({1, input.shape(0), input.shape(1), input.shape(2)});
    input = to4(input);
    onGpu({input}, "Interface::RateCvt");
    REQUIRE(time > 0, "Interface::RateCvt", "Input time can't be less than 1.");
    auto o = newSparse();
    check(spk_sparse_code(input.data(), ty(input.type()), input.shape(0), input.shape(1), input.shape(2),
                          input.shape(3), time, sort ? 1 : 0, out(o), nullptr),
          "Interface::RateCvt");
    return SparseTensor(o, input.device());
}

Tensor gather(SparseTensor input, Type type)
{
    const Shape s = input.shape();
    REQUIRE(s.size() == 5, "Interface::Gather", "Sparse can't be used without initialization.");
    return gather(input, Tensor(input.device(), type, {s[0], s[2], s[3], s[4]}));
}

Tensor gather(SparseTensor input, Tensor output)
{
    const spk_sparse &d = desc(input, "Interface::Gather");
    onGpu({output}, "Interface::Gather");
    REQUIRE(output.shape() == (Shape{d.n, d.c, d.h, d.w}), "Interface::Gather", "Tensor shapes are not compatible.");
    check(spk_sparse_gather(&d, output.data(), ty(output.type()), nullptr), "Interface::Gather");
    return output;
}

SparseTensor pool(SparseTensor input, Expand2 kernel, Expand2 stride, Expand4 pad)
{
    const spk_sparse &d = desc(input, "Interface::Pool");
    if (stride.get() == Shape{0, 0}) stride = kernel;
    checkItems("Kernel", kernel.get(), {1, 1}), checkItems("Stride", stride.get(), {1, 1});
    checkItems("Pad", pad.get(), {0, 0, 0, 0});
    auto o = newSparse();
    check(spk_sparse_pool(&d, kernel[0], kernel[1], stride[0], stride[1], pad[0], pad[1], pad[2], pad[3], out(o),
                          nullptr),
          "Interface::Pool");
    return SparseTensor(o, input.device());
}

SparseTensor inhibit(SparseTensor input)
{
    const spk_sparse &d = desc(input, "Interface::Inhibit");
    auto o = newSparse();
    check(spk_sparse_inhibit(&d, out(o), nullptr), "Interface::Inhibit");
    return SparseTensor(o, input.device());
}

Winners convwta(SparseTensor input, Expand2 radius, Size count)
{
    const spk_sparse &d = desc(input, "Interface::ConvWTA");
    checkItems("Radius", radius.get(), {0, 0});
    REQUIRE(count > 0, "Interface::ConvWTA", "Winners count can't be less than 1.");
    Tensor data(input.device(), Type::I64, {d.n, count, 5}), counts(input.device(), Type::I64, {d.n});
    check(spk_sparse_convwta(&d, radius[0], radius[1], count, data.data<spk_winner>(), counts.data<int64_t>(),
                             nullptr),
          "Interface::ConvWTA");
    return download(data, counts);
}

Winners fcwta(SparseTensor input, Size radius, Size count)
{
    const spk_sparse &d = desc(input, "Interface::FCWTA");
    REQUIRE(radius >= 0, "Interface::FCWTA", "Radius can't be less than 0.");
    REQUIRE(count > 0, "Interface::FCWTA", "Winners count can't be less than 1.");
    Tensor data(input.device(), Type::I64, {d.n, count, 5}), counts(input.device(), Type::I64, {d.n});
    check(spk_sparse_fcwta(&d, radius, count, data.data<spk_winner>(), counts.data<int64_t>(), nullptr),
          "Interface::FCWTA");
    return download(data, counts);
}
}  // namespace Sparse

SparseTensor Conv::operator()(SparseTensor input, F64 threshold)
{
    REQUIRE(_init, "Interface::Conv", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Conv", "Input device is not the same as layer's.");
    return Sparse::conv(input, kernel, threshold, _stride, _pad);
}

void Conv::stdp(SparseTensor input, const Winners &winners)
{
    REQUIRE(!stdpconfig.empty(), "Interface::STDP", "At least one STDP Config must be given.");
    REQUIRE(!winners.empty(), "Interface::STDP", "Winners can't be empty.");
    REQUIRE(_init, "Interface::STDP", "Layer is used before initializing.");
    REQUIRE(input.device() == _device, "Interface::Conv", "Input device is not the same as layer's.");
    const spk_sparse &d = desc(input, "Interface::STDP");
    onGpu({kernel}, "Interface::STDP");
    const DeviceWinners w = upload(winners, _device);
    const auto cfg = configs(stdpconfig);
    check(spk_sparse_convstdp(&d, kernel.data(), ty(kernel.type()), kernel.shape(0), kernel.shape(2),
                              kernel.shape(3), cfg.data(), Size(cfg.size()), w.data.data<spk_winner>(),
                              w.counts.data<int64_t>(), w.cap, _pad[0], _pad[1], _pad[2], _pad[3], nullptr),
          "Interface::STDP");
}

namespace
{
// The step as the reference runs it, one plain-path call per stage (layers
// outside the fused step's domain).
LayerStep plainStep(Conv &layer, Tensor input, F64 th, Expand2 radius, Size count, Expand2 pool, Expand2 poolStride,
                    bool learn)
{
    LayerStep r;
    Tensor pot = layer(input);
    threshold(pot, th);
    inhibit(pot);
    r.winners = convwta(pot, radius, count);
    r.output = fire(pot);
    if (learn) layer.stdp(input, r.winners, r.output);
    if (pool.get() != Shape{0, 0}) r.pooled = Spyker::pool(r.output, pool, poolStride);
    return r;
}
}  // namespace

LayerStep layerStep(Conv &layer, Tensor input, F64 threshold, Expand2 radius, Size count, Expand2 pool,
                    Expand2 poolStride, bool learn)
{
    input = to5(input);
    onGpu({input, layer.kernel}, "Interface::LayerStep");
    REQUIRE(input.type() == Type::U8, "Interface::LayerStep", "Input spikes must be u8.");
    REQUIRE(!learn || !layer.stdpconfig.empty(), "Interface::LayerStep", "At least one STDP Config must be given.");
    checkItems("Radius", radius.get(), {0, 0});
    REQUIRE(count > 0, "Interface::LayerStep", "Winners count can't be less than 1.");
    if (pool.get() != Shape{0, 0} && poolStride.get() == Shape{0, 0}) poolStride = pool;
    // the fused MFMA step (F32 master weights); a layer outside its domain
    // (kernel type, stride, step count, channels, K, threshold) runs on the
    // plain path in the kernel's type
    if (layer.kernel.type() != Type::F32 || layer.stride().get() != (Shape{1, 1}))
        return plainStep(layer, input, threshold, radius, count, pool, poolStride, learn);
    const Expand4 pad = layer.pad();
    const Shape os = convShape(input.shape(), layer.kernel.shape(), {1, 1}, pad.get());
    LayerStep r;
    r.output = Tensor(input.device(), Type::U8, os);
    if (pool.get() != Shape{0, 0})
    {
        checkItems("Kernel", pool.get(), {1, 1}), checkItems("Stride", poolStride.get(), {1, 1});
        r.pooled = Tensor(input.device(), Type::U8, poolShape(os, pool.get(), poolStride.get(), {0, 0, 0, 0}));
    }
    const Size n = input.shape(0);
    Tensor data(input.device(), Type::I64, {n, count, 5}), counts(input.device(), Type::I64, {n});
    const auto cfg = configs(layer.stdpconfig);
    const int32_t status = spk_stdp_layer_fused(
        input.data(), n, input.shape(1), input.shape(2), input.shape(3), input.shape(4), layer.kernel.data(),
        layer.kernel.shape(0), layer.kernel.shape(2), layer.kernel.shape(3), pad[0], pad[1], pad[2], pad[3], threshold,
        radius[0], radius[1], count, learn ? cfg.data() : nullptr, learn ? Size(cfg.size()) : 0, r.output.data(),
        r.pooled ? r.pooled.data() : nullptr, pool[0], pool[1], poolStride[0], poolStride[1],
        data.data<spk_winner>(), counts.data<int64_t>(), nullptr, nullptr);
    if (status == SPK_ERR_UNSUPPORTED)
        return plainStep(layer, input, threshold, radius, count, pool, poolStride, learn);
    check(status, "Interface::LayerStep");
    r.winners = download(data, counts);
    return r;
}
}  // namespace Spyker
