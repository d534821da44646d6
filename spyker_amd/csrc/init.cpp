// Host-side kernel generators, computed in F64 on the host as the reference
// does before casting and uploading (src/spyker/cpu/init.cpp:38-77,
// src/spyker/opers.cpp:702-717,739-754,838-849).
//
// The reference is built with FMA contraction; the products it contracts
// are written here as explicit std::fma so this file gives the same doubles
// whatever it is compiled with (it is compiled with -ffp-contract=off).
#include <cmath>
#include <mutex>
#include <random>

#include "common.hpp"

namespace
{
std::mutex g_rng_mu;
std::mt19937 g_rng(5489u);  // std::mt19937's default seed until spk_random_seed

void normalize(double *v, int64_t n)
{
    double sum = 0;
    for (int64_t i = 0; i < n; ++i) sum += v[i];
    for (int64_t i = 0; i < n; ++i) v[i] /= sum;
}
}  // namespace

using namespace spk;

extern "C" {

// Core::random_seed (src/spyker/impl.cpp:43)
int32_t spk_random_seed(int64_t seed)
{
    return guard([&] {
        std::lock_guard<std::mutex> lock(g_rng_mu);
        g_rng = std::mt19937(std::mt19937::result_type(seed));
    });
}

// normal_kernel (src/spyker/cpu/init.cpp:45-49): one fresh
// std::normal_distribution<double>(mean, std) per call, drawing from the
// shared std::mt19937.  The distribution is the Marsaglia polar method over
// 53-bit canonical uniforms (two 32-bit draws each); it is written out here
// so that the two sums the reference build contracts (x*x + y*y and
// z*std + mean) are single fused multiply-adds here too.
int32_t spk_normal_kernel(double *kernel, int64_t n, double mean, double std)
{
    return guard([&] {
        std::lock_guard<std::mutex> lock(g_rng_mu);
        auto uniform = [] {
            const double lo = double(g_rng() - std::mt19937::min());
            const double hi = double(g_rng() - std::mt19937::min());
            double r = (lo + hi * 4294967296.0) / 18446744073709551616.0;
            if (r >= 1.0) r = std::nextafter(1.0, 0.0);
            return r;
        };
        bool have = false;
        double saved = 0;
        for (int64_t i = 0; i < n; ++i)
        {
            double z;
            if (have)
                z = saved, have = false;
            else
            {
                double x, y, r2;
                do
                {
                    x = 2.0 * uniform() - 1.0;
                    y = 2.0 * uniform() - 1.0;
                    r2 = std::fma(x, x, y * y);
                } while (r2 > 1.0 || r2 == 0.0);
                const double mult = std::sqrt(-2 * std::log(r2) / r2);
                saved = x * mult, have = true;
                z = y * mult;
            }
            kernel[i] = std::fma(z, std, mean);
        }
    });
}

// gaussian_kernel (src/spyker/cpu/init.cpp:51-62)
int32_t spk_gaussian_kernel(double *kernel, int64_t k, double std)
{
    return guard([&] {
        SPK_ARG(k >= 1, "Kernel size can't be less than 1.");
        const int64_t size = k / 2;
        for (int64_t i = -size; i <= size; ++i)
            for (int64_t j = -size; j <= size; ++j)
            {
                const double r = std::sqrt(double(i * i + j * j)) / std;
                kernel[(i + size) * k + (j + size)] = std::exp(-.5 * (r * r));
            }
        normalize(kernel, k * k);
    });
}

// gabor_kernel (src/spyker/cpu/init.cpp:64-77); theta in radians.
int32_t spk_gabor_kernel(double *kernel, int64_t k, double sigma, double theta, double gamma, double lambda,
                         double psi)
{
    return guard([&] {
        SPK_ARG(k >= 1, "Kernel size can't be less than 1.");
        const double pi = 3.14159265358979323846;
        const int64_t size = k / 2;
        const double ct = std::cos(theta), st = std::sin(theta);
        for (int64_t i = -size; i <= size; ++i)
            for (int64_t j = -size; j <= size; ++j)
            {
                const double x = std::fma(double(j), st, double(i) * ct);
                const double y = std::fma(double(j), ct, double(-i) * st);
                const double a = std::exp(-std::fma(x, x, gamma * gamma * y * y) / (2 * sigma * sigma));
                const double b = std::cos(2 * pi * x / lambda + psi);
                kernel[(i + size) * k + (j + size)] = a * b;
            }
        normalize(kernel, k * k);
    });
}
}
