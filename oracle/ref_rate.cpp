// The reference's CPU rate coding driven from C, for pinning the oracle.
// TEST INFRASTRUCTURE ONLY: the product never links or loads it.
//
// src/spyker/cpu/code.cpp is compiled in this translation unit, from where it
// lies (the Makefile leaves it out of the per-file list), so that the
// per-sample steps it keeps internal -- CPU::latency, CPU::sorted and
// CPU::poisson (code.cpp:43-55, 94-101, 168-191) -- can be run with a table
// the caller gives.  The reference's own table comes from std::random_shuffle
// over std::rand inside an OpenMP loop, so its row order differs from run to
// run; its per-row contents are returned by ref_poisson_create.

#include <spyker/cpu/code.cpp>

#include <cstring>

using namespace Spyker;
using namespace Spyker::Core;

extern "C" {

// Core::poisson_create (cpu/init.cpp:120-149): rows x bins U16.
int ref_poisson_create(int64_t rows, int64_t bins, uint16_t *out)
{
    auto t = poisson_create(rows, bins);
    std::memcpy(out, t.data, sizeof(uint16_t) * size_t(rows * bins));
    CPU::deinit(t);
    return 0;
}

// CPU::poisson<I, O>(Vec2, Vec3) / CPU::poissort (code.cpp:192-230) for F32
// input and U8 output, with rnd standing for the shared generator's draw and
// `table` ([rows][1024]) for the cached Poisson table.  Returns -1 when a
// spike-time row falls outside the table.
int ref_rate_code_f32(const float *in, int64_t n, int64_t len, int64_t time, int32_t sort, uint64_t rnd_,
                      const uint16_t *table, int64_t rows, uint8_t *out)
{
    CPU::Poisson pois(1);
    auto own = pois.data;
    pois.data = Vec2<U16>((U16 *)table, rows, BINS);
    auto temp = CPU::init<U16>(len);
    auto index = CPU::init<U32>(len);
    auto times = CPU::init<U16>(len);
    auto values = CPU::init<U16>(len);
    std::mt19937::result_type rnd = std::mt19937::result_type(rnd_);
    int rc = 0;
    for (Size i = 0; i < n && rc == 0; ++i)
    {
        Vec1<F32> input((F32 *)in + i * len, len);
        Vec2<U8> output(out + i * time * len, time, len);
        CPU::Random random(rnd + i);
        if (sort)
            CPU::sorted(ARG(input), ARG(index), ARG(temp), time);
        else
            CPU::latency(ARG(input), ARG(temp), time);
        for (Size j = 0; j < len; ++j)
            if (temp(j) >= rows) rc = -1;
        if (rc == 0) CPU::poisson(ARG(temp), ARG(times), ARG(values), ARG(output), random, pois);
    }
    CPU::deinit(temp, index, times, values);
    pois.data = own;
    return rc;
}

// Core::cpu_rate_fire (cpu/fire.cpp:84-114, impl.h:299) for F32 input, U8 output.
int ref_rate_fire_f32(const float *in, int64_t n, int64_t t, int64_t len, float threshold, uint8_t *out)
{
    cpu_rate_fire(Dyn3((void *)in, Type::F32, n, t, len), Dyn3(out, Type::U8, n, t, len), Scalar(threshold));
    return 0;
}
}
