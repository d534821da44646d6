"""Layer-step throughput of the dense STDP layer on one MI355X.

Workload (the GPU-scale two-layer configuration): 1024 synthetic 28x28
stroke images, DoG on/off pairs (3x3 and 7x7, sigma 1/2 and 2/1) -> 4
channels, threshold 0.01, rank code T=15; layer 1 conv 4->30 5x5 pad 2,
fire at 15, pool 2x2 (inputs, built once, untimed).  The measured step is
layer 2 on [1024, 15, 30, 14, 14] spikes:

    conv 30->250 3x3 pad 1 -> threshold 10 -> inhibit -> convwta(r=1, k=8)
    -> fire -> STDP(a+=0.004, a-=-0.003) on the FP32 weights -> pool 3x3 s3

A step is one pass of that over the whole batch; value = images*timesteps/s
= 1024*15 / step time.  Two implementations run on the same inputs:

    plain  the FP32 per-stage kernels (bit-exact to the oracle)
    fused  the FP16-MFMA fused layer kernel (when built)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--path fused|plain]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, T_STEPS = 1024, 15
CIN, COUT, K, TH, RADIUS, COUNT = 30, 250, 3, 10.0, 1, 8
PEAK_FP32 = 157.3e12
PEAK_FP16 = 2.5e15
PEAK_HBM = 8.0e12


def make_inputs(dev, n=N, seed=2):
    import torch
    import spyker_amd as spk
    from spyker_amd import synthetic

    img = torch.from_numpy(synthetic.strokes(n, 28, seed=seed).astype(np.float32) / 255).to(dev)
    img = img.reshape(n, 1, 28, 28)
    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    filt = torch.cat([d3(img), d7(img)], dim=1).contiguous()  # [n, 4, 28, 28]
    spk.threshold(filt, 0.01)
    spikes = spk.code(filt, T_STEPS)  # [n, 15, 4, 28, 28]
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = np.clip(rng.normal(0.8, 0.05, (30, 4, 5, 5)), 0, 1).astype(np.float32)
    k2 = np.clip(rng.normal(0.8, 0.05, (COUT, CIN, K, K)), 0, 1).astype(np.float32)
    l1 = spk.pool(spk.fire(spk.conv(spikes, torch.from_numpy(k1).to(dev), 1, 2), 15.0), 2)
    torch.cuda.synchronize()
    return l1.contiguous(), k2


def plain_step(spk, x, conv):
    pot = conv(x)
    spk.threshold(pot, TH)
    spk.inhibit(pot)
    win = spk.convwta(pot, RADIUS, COUNT)
    out = spk.fire(pot)
    conv.stdp(x, win, out)
    return spk.pool(out, 3, 3), win


def cpu_baseline(x_host, k2, samples):
    """The oracle's C port of the same layer step on the host cores."""
    import oracle
    orc = oracle.Oracle()
    threads = os.cpu_count() or 1
    x = np.ascontiguousarray(x_host[:samples])
    n = x.shape[0]
    t0 = time.perf_counter()
    pot = orc.conv(x.reshape(n * T_STEPS, CIN, 14, 14), k2, 1, (1, 1, 1, 1), threads=threads)
    pot = pot.reshape(n, T_STEPS, COUT, 14, 14)
    pot = orc.threshold(pot, TH)
    pot = orc.inhibit(pot)
    w, c = orc.convwta(pot, RADIUS, COUNT)
    out = orc.fire(pot)
    orc.convstdp(x, k2, out, [(0.004, -0.003)], w, c, 1, (1, 1, 1, 1))
    orc.pool(out, 3, 3)
    dt = time.perf_counter() - t0
    return {"value": n * T_STEPS / dt, "unit": "images*timesteps/s", "cores": threads, "kind": "port",
            "sample": "%d images x %d steps of the same layer step (oracle C port, OpenMP)" % (n, T_STEPS)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", choices=["fused", "plain"], default="plain")
    ap.add_argument("--cpu-samples", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the layer path has no multi-GPU form in this build "
              "(--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    import spyker_amd as spk

    dev = torch.device("cuda", 0)
    x, k2 = make_inputs(dev)
    conv = spk.Conv(CIN, COUT, K, pad=1, device=dev)
    conv.kernel = torch.from_numpy(k2).to(dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]

    for _ in range(a.warmup):
        plain_step(spk, x, conv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        plain_step(spk, x, conv)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps

    # dominant kernel (the FP32 direct conv), HIP events on the launch stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = 5
    ev[0].record()
    for _ in range(reps):
        conv(x)
    ev[1].record()
    torch.cuda.synchronize()
    conv_s = ev[0].elapsed_time(ev[1]) / 1e3 / reps
    flop = 2.0 * N * T_STEPS * COUT * 14 * 14 * CIN * K * K
    achieved = flop / conv_s

    line = {
        "metric": "images*timesteps/s, conv+fire+wta+STDP layer, 1 GPU",
        "value": N * T_STEPS / dt,
        "unit": "images*timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "layer 2 of the two-layer stroke-image STDP network: batch 1024, 28x28 input, "
                               "T=15, conv 30->250 3x3 pad 1 on 14x14, th 10, convwta r1 k8, pool 3x3",
                   "path": "plain-fp32"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": None, "kernel": "conv_direct",
                     "kernel_ms": conv_s * 1e3},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(x.cpu().numpy(), k2, a.cpu_samples)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
