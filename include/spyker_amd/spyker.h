// C++ API of the MI355X backend: Spyker's public interface for the dense
// STDP layer path (src/spyker/opers.h, src/spyker/utils.h), with the same
// namespace, class and function names, argument meaning, defaults and
// errors, implemented over the C ABI of include/spyker_hip.h.
//
// Tensors live on the GPU (Kind::CUDA, i.e. HIP device memory) or on the
// host (Kind::CPU).  Every operation runs on the GPU; an operation given a
// host tensor throws instead of computing on the CPU.  Tensors made here own
// their memory (the library's caching allocator on the device); a tensor
// made from a caller's pointer with a no-op deleter wraps it, as the
// reference's pybind layer does for torch buffers (src/spyker/bind.cpp:184-187).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace Spyker
{
using Size = int64_t;
using F64 = double;
using Shape = std::vector<Size>;

/// Element types in the reference's order (src/spyker/utils.h:96-120).
enum class Type : int32_t
{
    I8,
    I16,
    I32,
    I64,
    U8,
    U16,
    U32,
    U64,
    F16,
    F32,
    F64,
};

/// Coding scheme (src/spyker/utils.h:122-128); this backend implements Rank.
enum class Code : int32_t
{
    Rank,
    Rate,
};

/// Device kind (src/spyker/utils.h:152-158); CUDA names the HIP GPU.
enum class Kind : int32_t
{
    CPU,
    CUDA,
};

/// Device to use for allocations and computations (src/spyker/utils.h:161-194).
class Device
{
    Kind _kind;
    Size _index;

public:
    Device(Kind kind = Kind::CPU, Size index = -1) : _kind(kind), _index(index) {}
    Kind kind() const { return _kind; }
    Size index() const { return _index; }
    bool operator==(const Device &o) const { return _kind == o._kind && (_index == o._index || _index < 0 || o._index < 0); }
    bool operator!=(const Device &o) const { return !(*this == o); }
    bool operator==(Kind kind) const { return _kind == kind; }
    bool operator!=(Kind kind) const { return _kind != kind; }
};

/// Bytes per element of a type.
Size typeSize(Type type);
/// Name of a type ("f32", ...).
std::string typeName(Type type);

/// N-dimensional contiguous row-major buffer on one device
/// (src/spyker/utils.h:202-330, the hot-path subset).
class Tensor
{
public:
    using Shared = std::shared_ptr<void>;

private:
    Shared _data;
    Size _offset = 0;
    Device _device;
    Type _type = Type::F32;
    Shape _shape;

    Tensor(Shared data, Size offset, Device device, Type type, Shape shape);

public:
    Tensor();
    /// New host tensor.
    Tensor(Type type, const Shape &shape);
    /// New tensor on `device`.
    Tensor(Device device, Type type, const Shape &shape);
    /// Wraps memory owned elsewhere (`data` carries the deleter).
    Tensor(Shared data, Device device, Type type, const Shape &shape);

    explicit operator bool() const { return bool(_data); }
    void *data() const;
    Size dims() const { return Size(_shape.size()); }
    Size numel() const;
    Size bytes() const { return numel() * typeSize(_type); }
    Shape shape() const { return _shape; }
    Size shape(Size index) const;
    Type type() const { return _type; }
    Device device() const { return _device; }

    /// Deep copy on the same device.
    Tensor copy() const;
    /// Conversion of the element type (a copy unless the type is the same).
    Tensor to(Type type) const;
    /// Transfer to another device (a copy unless it is the same device).
    Tensor to(Device device) const;
    /// Copies this tensor into `other` (same number of bytes).
    void to(Tensor other) const;
    /// Copies `tensor` into this one (same number of bytes).
    void from(Tensor tensor) const;
    /// View with another shape; one dimension may be -1.
    Tensor reshape(Shape shape) const;
    /// Sets every element to `value`.
    Tensor &fill(F64 value);
    /// View of sub-tensor `index` along the first dimension.
    Tensor operator[](Size index) const;

    template <typename T>
    T *data() const
    {
        return static_cast<T *>(data());
    }
};

/// Winner neuron (src/spyker/utils.h:483-499).
struct Winner
{
    Size c, t, z, y, x;
};

/// Per-sample winner lists (src/spyker/utils.h:502).
using Winners = std::vector<std::vector<Winner>>;

/// STDP parameters (src/spyker/utils.h:505-524).
struct STDPConfig
{
    F64 positive = 0, negative = 0;
    bool stabilize = true;
    F64 lower = 0, upper = 1;

    STDPConfig() {}
    STDPConfig(F64 positive, F64 negative, bool stabilize = true, F64 lower = 0, F64 upper = 1)
        : positive(positive), negative(negative), stabilize(stabilize), lower(lower), upper(upper)
    {
    }
};

/// Two-item size (src/spyker/opers.h:130-145): one value for both, or (y, x).
class Expand2
{
    Shape data;

public:
    Expand2() : data{0, 0} {}
    Expand2(Size _1) : data{_1, _1} {}
    Expand2(Size _1, Size _2) : data{_1, _2} {}
    Expand2(const Shape &shape);
    Size operator[](Size i) const { return data[i]; }
    Shape get() const { return data; }
};

/// Four-item size (src/spyker/opers.h:147-161): one value, (vertical,
/// horizontal) or (top, left, bottom, right).
class Expand4
{
    Shape data;

public:
    Expand4() : data{0, 0, 0, 0} {}
    Expand4(Size _1) : data{_1, _1, _1, _1} {}
    Expand4(Size _1, Size _2) : data{_1, _2, _1, _2} {}
    Expand4(Size _1, Size _2, Size _3, Size _4) : data{_1, _2, _3, _4} {}
    Expand4(const Shape &shape);
    Size operator[](Size i) const { return data[i]; }
    Shape get() const { return data; }
};

/// Difference of Gaussians parameters (src/spyker/opers.h:163-174).
struct DoGFilter
{
    F64 std1, std2;
    DoGFilter(F64 std1, F64 std2) : std1(std1), std2(std2) {}
};

/// Gabor parameters (src/spyker/opers.h:176-209); theta is given in degrees.
struct GaborFilter
{
    F64 sigma, theta, gamma, lambda, psi;
    GaborFilter(F64 sigma, F64 theta, F64 gamma, F64 lambda, F64 psi)
        : sigma(sigma), theta(theta * 0.017453292519943295), gamma(gamma), lambda(lambda), psi(psi)
    {
    }
};

// -- library control (src/spyker/opers.h:39-84) ---------------------------

void randomSeed(Size seed);
bool cudaAvailable();
Size cudaDeviceCount();
void cudaSetDevice(Size index);
Size cudaMemoryTaken(Size device = -1);
Size cudaMemoryUsed(Size device = -1);
void cudaCacheClear(Size device = -1);

// -- modules (src/spyker/opers.h:211-519) ---------------------------------

/// Difference of Gaussians filter bank.
class DoG
{
    bool _init = false;
    Device _device;
    Expand4 _pad;

public:
    Tensor kernel;

    DoG() {}
    DoG(Size size, const std::vector<DoGFilter> &filters, Expand4 pad = 0, Type type = Type::F32);
    DoG(Device device, Size size, const std::vector<DoGFilter> &filters, Expand4 pad = 0, Type type = Type::F32);
    Tensor operator()(Tensor input, Tensor output);
    Tensor operator()(Tensor input);
};

/// Gabor filter bank.
class Gabor
{
    bool _init = false;
    Device _device;
    Expand4 _pad;

public:
    Tensor kernel;

    Gabor() {}
    Gabor(Size size, const std::vector<GaborFilter> &filters, Expand4 pad = 0, Type type = Type::F32);
    Gabor(Device device, Size size, const std::vector<GaborFilter> &filters, Expand4 pad = 0,
          Type type = Type::F32);
    Tensor operator()(Tensor input, Tensor output);
    Tensor operator()(Tensor input);
};

/// Laplacian of Gaussian filter bank (src/spyker/opers.h:296-335): each std s
/// gives the Gaussians of s / sqrt 2 and s * sqrt 2; an input channel yields
/// [on, off] x count channels.
class LoG
{
    bool _init = false;
    Device _device;
    Expand4 _pad;

public:
    Tensor kernel;

    LoG() {}
    LoG(Size size, const std::vector<F64> &stds, Expand4 pad = 0, Type type = Type::F32);
    LoG(Device device, Size size, const std::vector<F64> &stds, Expand4 pad = 0, Type type = Type::F32);
    Tensor operator()(Tensor input, Tensor output);
    Tensor operator()(Tensor input);
    Device device() const { return _device; }
};

/// ZCA whitening (src/spyker/opers.h:339-375).  `fit` learns the F32 mean and
/// transform (FP64 inside) from the rows of an F32 device tensor; calling the
/// module whitens in place unless `inplace` is false.
class ZCA
{
public:
    Tensor mean;
    Tensor transform;

    ZCA() {}
    ZCA &fit(Tensor input, F64 epsilon, bool transform = false);
    Tensor operator()(Tensor input, bool inplace = true);
    static Tensor split(Tensor input, Tensor output);
    static Tensor split(Tensor input);
};

class SparseTensor;

/// 2D convolution layer with rank STDP.
class Conv
{
    bool _init = false;
    Device _device;
    Expand2 _stride;
    Expand4 _pad;

public:
    Tensor kernel;
    std::vector<STDPConfig> stdpconfig;

    Conv() {}
    Conv(Size input, Size output, Expand2 kernel, Expand2 stride = 1, Expand4 pad = 0, F64 mean = .5, F64 std = .02,
         Type type = Type::F32);
    Conv(Device device, Size input, Size output, Expand2 kernel, Expand2 stride = 1, Expand4 pad = 0, F64 mean = .5,
         F64 std = .02, Type type = Type::F32);
    Tensor operator()(Tensor input, Tensor output);
    Tensor operator()(Tensor input);
    void stdp(Tensor input, const Winners &winners, Tensor output);
    /// Sparse conv of spike lists (opers.cpp:866-871); the layer and the lists on the GPU.
    SparseTensor operator()(SparseTensor input, F64 threshold);
    /// STDP from spike lists (opers.cpp:882-889): potentiate where the input's last event
    /// is at a step <= the winner's, the sparse path's F32 rule.
    void stdp(SparseTensor input, const Winners &winners);
    Device device() const { return _device; }
    Expand2 stride() const { return _stride; }
    Expand4 pad() const { return _pad; }
};

/// Fully connected layer with rank STDP.
class FC
{
    bool _init = false;
    Device _device;

public:
    Tensor kernel;
    std::vector<STDPConfig> stdpconfig;

    FC() {}
    FC(Size input, Size output, F64 mean = .5, F64 std = .02, Type type = Type::F32);
    FC(Device device, Size input, Size output, F64 mean = .5, F64 std = .02, Type type = Type::F32);
    Tensor operator()(Tensor input, Tensor output);
    Tensor operator()(Tensor input);
    void stdp(Tensor input, const Winners &winners, Tensor output);
    Device device() const { return _device; }
};

// -- functions (src/spyker/opers.h:527-753) -------------------------------

Tensor dog(Tensor input, Tensor kernel, Expand4 pad = 0);
Tensor dog(Tensor input, Tensor kernel, Tensor output, Expand4 pad = 0);
Tensor gabor(Tensor input, Tensor kernel, Expand4 pad = 0);
Tensor gabor(Tensor input, Tensor kernel, Tensor output, Expand4 pad = 0);
Tensor conv(Tensor input, Tensor kernel, Expand2 stride = 1, Expand4 pad = 0);
Tensor conv(Tensor input, Tensor kernel, Tensor output, Expand2 stride = 1, Expand4 pad = 0);
Tensor fc(Tensor input, Tensor kernel);
Tensor fc(Tensor input, Tensor kernel, Tensor output);
Tensor pad(Tensor input, Expand4 pad, F64 value = 0);
Tensor pad(Tensor input, Tensor output, Expand4 pad, F64 value = 0);
Tensor threshold(Tensor input, F64 threshold, F64 value = 0, bool inplace = true);
Tensor code(Tensor input, Size time, bool sort = true, Type type = Type::U8, Code code = Code::Rank);
Tensor code(Tensor input, Tensor output, Size time, bool sort = true, Code code = Code::Rank);
Tensor fire(Tensor input, F64 threshold = 0, Type type = Type::U8, Code code = Code::Rank);
Tensor fire(Tensor input, Tensor output, F64 threshold = 0, Code code = Code::Rank);
Tensor pool(Tensor input, Expand2 kernel, Expand2 stride = 0, Expand4 pad = 0);
Tensor pool(Tensor input, Tensor output, Expand2 kernel, Expand2 stride = 0, Expand4 pad = 0);
Tensor inhibit(Tensor input, F64 threshold = 0, bool inplace = true);
Winners fcwta(Tensor input, Size radius, Size count, F64 threshold = 0);
Winners convwta(Tensor input, Expand2 radius, Size count, F64 threshold = 0);

// -- readout (opers.h:662-699) ----------------------------------------------

/// First spike step of every element (Code::Rank), or the last step (Code::Rate):
/// [N, T, ...] -> [N, ...].
Tensor gather(Tensor input, F64 threshold = 0, Type type = Type::U8, Code code = Code::Rank);
Tensor gather(Tensor input, Tensor output, F64 threshold = 0, Code code = Code::Rank);
/// Spike trains from first-spike times: [N, ...] -> [N, time, ...].
Tensor scatter(Tensor input, Size time, Type type = Type::U8);
Tensor scatter(Tensor input, Tensor output);

// -- sparse event path (utils.h:369-420, opers.h:763-828) ----------------------

/// Spike lists of a [N, T, C, H, W] tensor on the GPU: per sample and step the
/// spikes that are new at that step (base.h:493-528), C, H, W <= 255.
class SparseTensor
{
    std::shared_ptr<void> _data;  // the spk_sparse descriptor; its lists are released with it
    Device _device;

public:
    SparseTensor();
    /// Events at each element's first step with a value > threshold (U8 or F32 input).
    explicit SparseTensor(Tensor tensor, F64 threshold = 0);
    /// Adopts a descriptor filled by the C ABI (the Sparse functions use it).
    SparseTensor(std::shared_ptr<void> descriptor, Device device);

    explicit operator bool() const { return bool(_data); }
    void *data() const { return _data.get(); }
    Size dims() const;
    /// Number of events.
    Size numel() const;
    Size bytes() const;
    Shape shape() const;
    Size shape(Size index) const;
    F64 sparsity() const;
    Device device() const { return _device; }
    SparseTensor copy() const;
    /// Cumulative U8 spikes [N, T, C, H, W].
    Tensor dense() const;
};

namespace Sparse
{
SparseTensor conv(SparseTensor input, Tensor kernel, F64 threshold, Expand2 stride = 1, Expand4 pad = 0);
SparseTensor pad(SparseTensor input, Expand4 pad);
SparseTensor code(Tensor input, Size time, bool sort = true);
Tensor gather(SparseTensor input, Type type = Type::U8);
Tensor gather(SparseTensor input, Tensor output);
SparseTensor pool(SparseTensor input, Expand2 kernel, Expand2 stride = 0, Expand4 pad = 0);
SparseTensor inhibit(SparseTensor input);
Winners convwta(SparseTensor input, Expand2 radius, Size count);
/// No single reference call: the winners of fcwta(input.dense()), as conv
/// coordinates with their spike step (Winner{0, t, z, y, x}).
Winners fcwta(SparseTensor input, Size radius, Size count);
}  // namespace Sparse

// -- fused layer step (no single counterpart in the reference) -------------

/// Output of layerStep: the layer's spikes, its winners and (if a pool
/// kernel was given) the pooled spikes.
struct LayerStep
{
    Tensor output;
    Winners winners;
    Tensor pooled;
};

/// One training step of `layer` on U8 spikes, as the reference runs it in
/// separate calls (examples/kheradpisheh_mnist_dense.py:62-64):
///     pot = inhibit(threshold(layer(input), threshold)); winners = convwta(pot, radius, count);
///     output = fire(pot); layer.stdp(input, winners, output); pooled = pool(output, pool, poolStride)
/// on FP16 MFMA with FP32 accumulation and the FP32 master weights in
/// `layer.kernel`.  `learn = false` leaves the weights unchanged.  A pool
/// kernel of 0 skips pooling.  A layer outside the fused step's domain
/// (stride other than 1, more than 32 steps or 256 output channels, Cin*Kh*Kw
/// above 512, a negative threshold) runs the same calls on the FP32 plain path.
LayerStep layerStep(Conv &layer, Tensor input, F64 threshold, Expand2 radius, Size count, Expand2 pool = 0,
                    Expand2 poolStride = 0, bool learn = true);
}  // namespace Spyker
