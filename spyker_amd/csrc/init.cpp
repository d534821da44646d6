// Host-side kernel generators, computed in F64 on the host as the reference
// does before casting and uploading (src/spyker/cpu/init.cpp:38-77,
// src/spyker/opers.cpp:702-717,739-754,838-849).
//
// The reference is built with FMA contraction; the products it contracts
// are written here as explicit std::fma so this file gives the same doubles
// whatever it is compiled with (it is compiled with -ffp-contract=off).
#include <algorithm>
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

namespace spk
{
// Generator() of the reference (src/spyker/base.h:577): one draw of the shared
// std::mt19937, e.g. the per-call seed base of the rate code (cpu/code.cpp:195).
uint64_t generator_draw()
{
    std::lock_guard<std::mutex> lock(g_rng_mu);
    return uint64_t(g_rng());
}

// Core::poisson_create (src/spyker/cpu/init.cpp:120-149).  Row i is the
// interval table of lambda = i + 1: round(ztp(j, lambda) * bins) copies of
// every interval j in [max(1, lambda - range), lambda + range], range =
// ceil(scale * sqrt(lambda)) with scale 5 in the first row and 4 after, the
// rest of the row lambda; ztp is the zero-truncated Poisson probability
// lambda^j / (j! (e^lambda - 1)) in F32, accumulated as there.  The reference
// then permutes each row with std::random_shuffle, whose std::rand state
// (shared, and touched from an OpenMP loop) differs from run to run; here each
// row is permuted by a Fisher-Yates pass over std::mt19937(5489 + row), the
// same in every process.  A draw picks a uniform position of its row, so the
// permutation leaves the code's distribution unchanged.
void poisson_rows(uint16_t *table, int64_t rows, int64_t bins)
{
    auto ztp = [](uint16_t k, float a) {
        float p = 1 / (std::exp(a) - 1);
        for (int64_t i = 1; i <= k; ++i) p *= a / float(i);
        return p;
    };
    for (int64_t i = 0; i < rows; ++i)
    {
        const float scale = i < 1 ? 5 : 4;
        const uint16_t lambda = uint16_t(i + 1);
        const uint16_t range = uint16_t(std::ceil(scale * std::sqrt(double(lambda))));
        const uint16_t start = uint16_t(std::max(1, int(lambda) - int(range)));
        const uint16_t end = uint16_t(lambda + range);
        uint16_t *row = table + i * bins;
        int64_t index = 0;
        for (uint16_t j = start; j <= end; ++j)
        {
            int64_t count = int64_t(std::nearbyint(ztp(j, float(lambda)) * float(bins)));
            if (index + count > bins) count = bins - index;
            std::fill(row + index, row + index + count, j);
            index += count;
        }
        std::fill(row + index, row + bins, lambda);
        std::mt19937 shuffle(uint32_t(5489u + uint64_t(i)));
        for (int64_t j = bins - 1; j > 0; --j) std::swap(row[j], row[shuffle() % uint64_t(j + 1)]);
    }
}
}  // namespace spk

using namespace spk;

extern "C" {

// Core::poisson_create (src/spyker/impl.h:190): [rows][bins] U16 on the host.
int32_t spk_poisson_table(uint16_t *table, int64_t rows, int64_t bins)
{
    return guard([&] {
        SPK_ARG(rows >= 1 && rows <= 65535, "Poisson table: rows must be in [1, 65535].");
        SPK_ARG(bins >= 1, "Poisson table: bins can't be less than 1.");
        poisson_rows(table, rows, bins);
    });
}

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
