// The reference's CUDA backend seam (Core::cuda_*, src/spyker/impl.h:68-322)
// implemented over this library's C ABI (include/spyker_hip.h).
//
// A maintainer compiles this file into Spyker in place of src/spyker/cuda/*.cu,
// with the reference's own SPYKER_USE_CUDA definition (CMakeLists.txt:167,240)
// and its headers on the include path, and links -lspyker_hip.  Nothing above
// the seam changes: opers.cpp, utils.cpp, bind.cpp and the Python package
// still route Kind::CUDA tensors through Choose(...) (impl.h:45-52), which now
// lands here.  torch-ROCm "cuda" tensors are HIP device memory, so the
// zero-copy wrap of bind.cpp:184-187 works unchanged.
//
// Every call is ordered on the null stream, as the reference's backend is
// (it launches on the legacy default stream and copies with blocking
// cudaMemcpy); calls that return host data (winners) wait for their copies.
// Functions off the dense layer path throw the reference's error form.
#include <spyker/impl.h>
#include <spyker_hip.h>

#include <algorithm>
#include <iostream>
#include <vector>

namespace Spyker
{
namespace Core
{
namespace
{
// A failing call throws as SpykerAssert does (utils.h:50-66), with the
// library's message.
void check(int32_t status, const char *scope) { SpykerAssert(status == SPK_OK, scope, spk_last_error()); }

[[noreturn]] void absent(const char *scope)
{
    SpykerAssert(false, scope, "This function is not implemented by the HIP backend.");
    throw std::runtime_error(scope);  // not reached
}

int32_t ty(Type t) { return int32_t(t); }  // spk_type follows the order of Type (utils.h:96-117)

bool cache_on = true;

// Host winners <-> device records [n][maxcount] plus counts.
struct DeviceWinners
{
    void *rec = nullptr, *cnt = nullptr;
    Size maxcount = 1;
    DeviceWinners(const Winners &winners, Size n, const char *scope)
    {
        for (auto &w : winners) maxcount = std::max(maxcount, Size(w.size()));
        std::vector<spk_winner> hw(size_t(n * maxcount), spk_winner{-1, 0, 0, 0, 0});
        std::vector<int64_t> hc(size_t(n), 0);
        for (Size i = 0; i < n && i < Size(winners.size()); ++i)
        {
            hc[size_t(i)] = Size(winners[size_t(i)].size());
            for (Size k = 0; k < hc[size_t(i)]; ++k)
            {
                const Winner &e = winners[size_t(i)][size_t(k)];
                hw[size_t(i * maxcount + k)] = spk_winner{e.c, e.t, e.z, e.y, e.x};
            }
        }
        check(spk_alloc(&rec, Size(hw.size() * sizeof(spk_winner))), scope);
        check(spk_alloc(&cnt, Size(hc.size() * sizeof(int64_t))), scope);
        check(spk_copy(hw.data(), rec, Size(hw.size() * sizeof(spk_winner)), nullptr), scope);
        check(spk_copy(hc.data(), cnt, Size(hc.size() * sizeof(int64_t)), nullptr), scope);
    }
    ~DeviceWinners()
    {
        // the call that read them is ordered before any later use of the blocks
        spk_dealloc(rec), spk_dealloc(cnt);
    }
};

Winners download(void *rec, void *cnt, Size n, Size count, const char *scope)
{
    std::vector<spk_winner> hw(static_cast<size_t>(n * count));
    std::vector<int64_t> hc(static_cast<size_t>(n));
    check(spk_copy(rec, hw.data(), Size(hw.size() * sizeof(spk_winner)), nullptr), scope);
    check(spk_copy(cnt, hc.data(), Size(hc.size() * sizeof(int64_t)), nullptr), scope);
    check(spk_synchronize(nullptr), scope);
    Winners out(static_cast<size_t>(n));
    for (Size i = 0; i < n; ++i)
        for (Size k = 0; k < hc[size_t(i)]; ++k)
        {
            const spk_winner &e = hw[size_t(i * count + k)];
            out[size_t(i)].push_back(Winner{e.c, e.t, e.z, e.y, e.x});
        }
    return out;
}

std::vector<spk_stdp_config> configs(const std::vector<STDPConfig> &config)
{
    std::vector<spk_stdp_config> out;
    for (auto &c : config)
        out.push_back(spk_stdp_config{F64(c.positive), F64(c.negative), c.stabilize ? 1 : 0, F64(c.lower),
                                      F64(c.upper)});
    return out;
}
}  // namespace

// -- devices and memory (impl.h:70-126) ----------------------------------------

bool cuda_available()
{
    int32_t count = 0;
    return spk_device_count(&count) == SPK_OK && count > 0;
}

Size cuda_device_count()
{
    int32_t count = 0;
    check(spk_device_count(&count), "Core::CUDA");
    return count;
}

void cuda_set_device(Size device) { check(spk_set_device(int32_t(device)), "Core::CUDA"); }

Size cuda_current_device()
{
    int32_t device = 0;
    check(spk_current_device(&device), "Core::CUDA");
    return device;
}

// The library is built for one target, gfx950 (major 9, minor 5).
float cuda_current_arch() { return 9.5f; }

std::vector<Size> cuda_arch_list() { return {95}; }

std::vector<Size> cuda_device_arch() { return std::vector<Size>(size_t(cuda_device_count()), 95); }

Size cuda_memory_total()
{
    int64_t total = 0;
    check(spk_memory_info(nullptr, &total), "Core::CUDA");
    return total;
}

Size cuda_memory_free()
{
    int64_t free = 0;
    check(spk_memory_info(&free, nullptr), "Core::CUDA");
    return free;
}

Size cuda_memory_taken(Size)
{
    int64_t bytes = 0;
    check(spk_memory_taken(&bytes), "Core::CUDA");
    return bytes;
}

Size cuda_memory_used(Size)
{
    int64_t bytes = 0;
    check(spk_memory_used(&bytes), "Core::CUDA");
    return bytes;
}

bool cuda_cache_enabled(Size) { return cache_on; }

// With the cache off, every free returns the cached blocks to the driver.
void cuda_cache_enable(bool enabled, Size) { cache_on = enabled; }

void cuda_cache_clear(Size) { check(spk_cache_clear(), "Core::CUDA"); }

void cuda_cache_print(Size device)
{
    std::cout << "HIP cache: " << cuda_memory_taken(device) << " bytes taken, " << cuda_memory_used(device)
              << " bytes used" << std::endl;
}

void *pinned_alloc(Size) { absent("Core::CUDA pinned_alloc"); }
void pinned_dealloc(void *) { absent("Core::CUDA pinned_dealloc"); }
void *unified_alloc(Size) { absent("Core::CUDA unified_alloc"); }
void unified_dealloc(void *) { absent("Core::CUDA unified_dealloc"); }

void cpu2cuda(Size size, void *input, void *output)
{
    check(spk_copy(input, output, size, nullptr), "Core::CUDA");
    check(spk_synchronize(nullptr), "Core::CUDA");
}

void cuda2cpu(Size size, void *input, void *output)
{
    check(spk_copy(input, output, size, nullptr), "Core::CUDA");
    check(spk_synchronize(nullptr), "Core::CUDA");
}

void cuda_conv_clear() {}  // no per-shape conv state to drop

void cuda_poisson_clear() { check(spk_poisson_clear(), "Core::CUDA"); }

void cuda_conv_options(Size, Size, Size) {}  // one conv algorithm per path: nothing to choose

void *cuda_alloc(Size size)
{
    void *p = nullptr;
    check(spk_alloc(&p, size), "Core::CUDA");
    return p;
}

void cuda_dealloc(void *data)
{
    check(spk_dealloc(data), "Core::CUDA");
    if (!cache_on) check(spk_cache_clear(), "Core::CUDA");
}

void cuda_copy(Size size, void *input, void *output)
{
    check(spk_copy(input, output, size, nullptr), "Core::CUDA");
    check(spk_synchronize(nullptr), "Core::CUDA");
}

void cuda_cast(Size size, Dyn input, Dyn output)
{
    check(spk_cast(input.data, ty(input.type), output.data, ty(output.type), size, nullptr), "Core::Cast");
}

void cuda_fill(Size size, Dyn input, Scalar value)
{
    check(spk_fill(input.data, ty(input.type), size, F64(value), nullptr), "Core::Fill");
}

// -- filters (impl.h:196-222) -------------------------------------------------

void cuda_dog(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad)
{
    // input [planes][1][h][w], kernel [2*count][1][k][k], output [planes][count][ho][wo] (opers.cpp:255-258)
    check(spk_dog(input.data, kernel.data, output.data, input.t, input.y, input.x, output.z, kernel.y, pad.t, pad.z,
                  pad.y, pad.x, nullptr),
          "Core::DoG");
}

void cuda_gabor(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad)
{
    check(spk_gabor(input.data, kernel.data, output.data, input.t, input.y, input.x, output.z, kernel.y, pad.t, pad.z,
                    pad.y, pad.x, nullptr),
          "Core::Gabor");
}

void cuda_log(Dyn4 input, Dyn4 kernel, Dyn4 output, Len4 pad)
{
    check(spk_log(input.data, kernel.data, output.data, input.t, input.y, input.x, output.z, kernel.y, pad.t, pad.z,
                  pad.y, pad.x, nullptr),
          "Core::LoG");
}

void cuda_zca_fit(Dyn2 input, Dyn1 mean, Dyn2 trans, Scalar epsilon, bool transform)
{
    check(spk_zca_fit(input.data, mean.data, trans.data, input.y, input.x, F64(epsilon), transform ? 1 : 0, nullptr),
          "Core::ZCA");
}

void cuda_zca_trans(Dyn2 input, Dyn1 mean, Dyn2 trans)
{
    check(spk_zca_trans(input.data, mean.data, trans.data, input.y, input.x, nullptr), "Core::ZCA");
}

void cuda_zca_split(Dyn2 input, Dyn3 output)
{
    check(spk_zca_split(input.data, output.data, ty(input.type), input.y, input.x, nullptr), "Core::ZCA");
}

// -- contractions (impl.h:224-234) --------------------------------------------

void cuda_conv(Dyn4 input, Dyn4 kernel, Dyn4 output, Len2 stride, Len4 pad)
{
    check(spk_conv(input.data, ty(input.type), kernel.data, ty(kernel.type), output.data, ty(output.type), input.t,
                   input.z, input.y, input.x, kernel.t, kernel.y, kernel.x, stride.y, stride.x, pad.t, pad.z, pad.y,
                   pad.x, nullptr),
          "Core::Conv");
}

void cuda_fc(Dyn3 input, Dyn2 kernel, Dyn3 output)
{
    SpykerAssert(kernel.type == Type::F32 && output.type == Type::F32, "Core::FC",
                 "The HIP backend runs F32 kernels.");
    check(spk_fc(input.data, ty(input.type), kernel.data, output.data, input.z * input.y, input.x, kernel.y, nullptr),
          "Core::FC");
}

void cuda_signfc(Dyn3, Dyn2, Dyn3) { absent("Core::SignFC"); }

// -- elementwise, code, pool, inhibit, winners (impl.h:236-305) ---------------

void cuda_pad(Dyn3 input, Dyn3 output, Len4 pad, Scalar value)
{
    check(spk_pad(input.data, output.data, ty(input.type), input.z, input.y, input.x, pad.t, pad.z, pad.y, pad.x,
                  F64(value), nullptr),
          "Core::Pad");
}

void cuda_threshold(Dyn1 input, Scalar threshold, Scalar value)
{
    check(spk_threshold(input.data, ty(input.type), input.x, F64(threshold), F64(value), nullptr), "Core::Threshold");
}

void cuda_quantize(Dyn1, Scalar, Scalar, Scalar) { absent("Core::Quantize"); }

void cuda_rank_code(Dyn2 input, Dyn3 output, bool sort)
{
    check(spk_rank_code(input.data, ty(input.type), output.data, ty(output.type), input.y, input.x, output.y,
                        sort ? 1 : 0, nullptr),
          "Core::RankCode");
}

void cuda_rank_infinite(Dyn3, Scalar) { absent("Core::RankInfinite"); }

void cuda_rank_fire(Dyn1 input, Dyn1 output, Scalar threshold)
{
    check(spk_rank_fire(input.data, ty(input.type), output.data, ty(output.type), input.x, F64(threshold), nullptr),
          "Core::RankFire");
}

void cuda_rank_gather(Dyn3 input, Dyn2 output, Scalar threshold)
{
    check(spk_rank_gather(input.data, ty(input.type), output.data, ty(output.type), input.z, input.y, input.x,
                          F64(threshold), nullptr),
          "Core::RankGather");
}

void cuda_rank_scatter(Dyn2 input, Dyn3 output)
{
    check(spk_rank_scatter(input.data, ty(input.type), output.data, ty(output.type), input.y, input.x, output.y,
                           nullptr),
          "Core::RankScatter");
}

void cuda_rank_pool(Dyn4 input, Dyn4 output, Len2 kernel, Len2 stride, Len4 pad)
{
    check(spk_rank_pool(input.data, output.data, ty(input.type), input.t * input.z, input.y, input.x, kernel.y,
                        kernel.x, stride.y, stride.x, pad.t, pad.z, pad.y, pad.x, nullptr),
          "Core::RankPool");
}

void cuda_rank_inhibit(Dyn4 input, Scalar threshold)
{
    check(spk_rank_inhibit(input.data, ty(input.type), input.t, input.z, input.y, input.x, F64(threshold), nullptr),
          "Core::RankInhibit");
}

Winners cuda_rank_fcwta(Dyn3 input, Size radius, Size count, Scalar threshold)
{
    const Size n = input.z;
    void *rec = nullptr, *cnt = nullptr;
    check(spk_alloc(&rec, n * count * Size(sizeof(spk_winner))), "Core::FCWTA");
    check(spk_alloc(&cnt, n * Size(sizeof(int64_t))), "Core::FCWTA");
    const int32_t status = spk_rank_fcwta(input.data, ty(input.type), n, input.y, input.x, radius, count,
                                          F64(threshold), (spk_winner *)rec, (int64_t *)cnt, nullptr);
    Winners out = status == SPK_OK ? download(rec, cnt, n, count, "Core::FCWTA") : Winners();
    spk_dealloc(rec), spk_dealloc(cnt);
    check(status, "Core::FCWTA");
    return out;
}

Winners cuda_rank_convwta(Dyn5 input, Len2 radius, Size count, Scalar threshold)
{
    const Size n = input.u;
    void *rec = nullptr, *cnt = nullptr;
    check(spk_alloc(&rec, n * count * Size(sizeof(spk_winner))), "Core::ConvWTA");
    check(spk_alloc(&cnt, n * Size(sizeof(int64_t))), "Core::ConvWTA");
    const int32_t status =
        spk_rank_convwta(input.data, ty(input.type), n, input.t, input.z, input.y, input.x, radius.y, radius.x, count,
                         F64(threshold), (spk_winner *)rec, (int64_t *)cnt, nullptr);
    Winners out = status == SPK_OK ? download(rec, cnt, n, count, "Core::ConvWTA") : Winners();
    spk_dealloc(rec), spk_dealloc(cnt);
    check(status, "Core::ConvWTA");
    return out;
}

void cuda_rank_convstdp(Dyn5 input, Dyn4 kernel, Dyn5 output, const std::vector<STDPConfig> &config,
                        const Winners &winners, Len2 stride, Len4 pad)
{
    const DeviceWinners w(winners, input.u, "Core::STDP");
    const auto cfg = configs(config);
    check(spk_rank_convstdp(input.data, kernel.data, ty(kernel.type), output.data, input.u, input.t, input.z, input.y,
                            input.x, kernel.t, kernel.y, kernel.x, output.y, output.x, cfg.data(), Size(cfg.size()),
                            (const spk_winner *)w.rec, (const int64_t *)w.cnt, w.maxcount, stride.y, stride.x, pad.t,
                            pad.z, pad.y, pad.x, nullptr),
          "Core::STDP");
}

void cuda_rank_fcstdp(Dyn3 input, Dyn2 kernel, Dyn3 output, const std::vector<STDPConfig> &config,
                      const Winners &winners)
{
    const DeviceWinners w(winners, input.z, "Core::STDP");
    const auto cfg = configs(config);
    check(spk_rank_fcstdp(input.data, kernel.data, ty(kernel.type), output.data, input.z, input.y, input.x, kernel.y,
                          cfg.data(), Size(cfg.size()), (const spk_winner *)w.rec, (const int64_t *)w.cnt, w.maxcount,
                          nullptr),
          "Core::STDP");
}

// -- rate code and the functions off the layer path (impl.h:295-322) ----------

void cuda_rate_code(Dyn2 input, Dyn3 output, bool sort)
{
    check(spk_rate_code(input.data, ty(input.type), output.data, ty(output.type), input.y, input.x, output.y,
                        sort ? 1 : 0, nullptr),
          "Core::RateCode");
}

void cuda_rate_fire(Dyn3 input, Dyn3 output, Scalar threshold)
{
    check(spk_rate_fire(input.data, ty(input.type), output.data, ty(output.type), input.z, input.y, input.x,
                        F64(threshold), nullptr),
          "Core::RateFire");
}

void cuda_rate_gather(Dyn3 input, Dyn2 output, Scalar threshold)
{
    (void)threshold;  // unused by the reference too: the last step's count (cpu/time.cpp:88-92)
    check(spk_rate_gather(input.data, ty(input.type), output.data, ty(output.type), input.z, input.y, input.x,
                          nullptr),
          "Core::RateGather");
}

void cuda_rate_pool(Dyn5, Dyn4, Dyn5, Len2, Len2, Len4) { absent("Core::RatePool"); }
void cuda_canny(Dyn3, Dyn3, Scalar, Scalar) { absent("Core::Canny"); }
void cuda_backward(Dyn2, Dyn2, Dyn1, Size, Scalar) { absent("Core::Backward"); }
void cuda_labelize(Dyn3, Dyn1, Scalar) { absent("Core::Labelize"); }
void cuda_fcbackward(Dyn2, Dyn2, Dyn2, Dyn2, Dyn2, BPConfig &) { absent("Core::FCBackward"); }
}  // namespace Core
}  // namespace Spyker
