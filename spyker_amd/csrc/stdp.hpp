// Rank STDP launch interface shared by the plain entry points and the fused
// layer step.
#pragma once

#include "common.hpp"

namespace spk
{
struct StdpGeom
{
    // FC uses cin = input features, cout = output neurons, the rest unused.
    int32_t t, cin, h, w, cout, kh, kw, ho, wo, sy, sx, pt, pl;
    int64_t n, maxcount;
    int32_t ncfg;
};

struct StdpConfigs
{
    spk_stdp_config c[SPK_MAX_CONFIGS];
    int32_t n;
};

// One update of one weight (stdp.cpp:121-137): the F64 formula with the
// multiply-add the reference's build contracts `old + rate * value` into,
// then the clamp, rounded once to the weight's type (F32, F16 or F64).
// The same for an F16 or F64 weight (stdp.cpp:121-137 with K = F16/F64):
// the F64 result is rounded once to K.
template <typename K>
__device__ __forceinline__ K stdp_update(K w, bool potentiate, const spk_stdp_config &cfg)
{
    const double old = double(w);
    double value = cfg.stabilize ? (old - cfg.lower) * (cfg.upper - old) : 1.0;
    value = __builtin_fma(potentiate ? cfg.positive : cfg.negative, value, old);
    const double hi = value < cfg.upper ? value : cfg.upper;
    const double r = cfg.lower < hi ? hi : cfg.lower;
    return K(r);
}

// The sparse path's update of one weight (src/spyker/sparse/stdp.cpp:111-118):
// F32 weight arithmetic with the F64 terms of the configuration,
//     mult = (F32)((k - lower) * (upper - k))     (1 without stabilize)
//     k = (F32)(k + rate * mult)                  (contracted to one F64 fma)
//     k = (F32)max(lower, min(k, upper))
__device__ __forceinline__ float sparse_stdp_update(float k, bool potentiate, const spk_stdp_config &cfg)
{
    const float mult = cfg.stabilize ? float((double(k) - cfg.lower) * (cfg.upper - double(k))) : 1.0f;
    const float r = float(__builtin_fma(potentiate ? cfg.positive : cfg.negative, double(mult), double(k)));
    const double hi = double(r) < cfg.upper ? double(r) : cfg.upper;
    return float(cfg.lower < hi ? hi : cfg.lower);
}

StdpConfigs pack_configs(const spk_stdp_config *configs, int64_t nconfigs);

// Bytes of the per-winner bit records the launch needs as workspace.
int64_t stdp_bits_bytes(const StdpGeom &g, bool fc);

// Enqueues the record and apply kernels.  flag must be a device int32 that
// is zero on entry; a nonzero value afterwards means the call was rejected
// (1: configuration index out of range, 2: winner index out of range) and no
// weight changed.
void stdp_launch(const uint8_t *in, void *kernel, int32_t ktype, const uint8_t *out, const StdpGeom &g,
                 const StdpConfigs &cfg, const spk_winner *winners, const int64_t *counts, uint32_t *bits,
                 int32_t *flag, bool fc, hipStream_t s);
// The apply launch alone, for records made elsewhere (the sparse path):
// slot[e] = {output channel or -1, config}, bits [slots][ceil(K/32)].
// ktype: SPK_F32, SPK_F16 or SPK_F64 (the sparse rule takes F32 only).
void stdp_apply(void *kernel, int32_t ktype, const StdpGeom &g, const StdpConfigs &cfg, const int2 *slot,
                const uint32_t *bits, const int32_t *flag, bool fc, bool sparse_rule, hipStream_t s);
}  // namespace spk
