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

    fused  the FP16-MFMA fused layer step (spk_stdp_layer_fused) -> `value`
    plain  the FP32 per-stage kernels, bit-exact to the oracle -> `plain`

A second, smaller line item (`sparse`) times the sparse event-stream
workload: 256 DVS-like samples of 2x128x128 events over 20 bins (2% density
modulated by moving blobs) -> sparse lists -> sparse conv 2->64 5x5 pad 2,
threshold 8 -> fcwta k16 on the densified output -> sparse STDP of the
winners, in samples*bins/s, next to the reference's own sparse CPU code.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, T_STEPS = 1024, 15
CIN, COUT, K, H, TH, RADIUS, COUNT = 30, 250, 3, 14, 10.0, 1, 8
FLOP_PER_IT = 2.0 * COUT * H * H * CIN * K * K          # dense time-unrolled conv, per image*timestep
BYTES_PER_IT = CIN * H * H + COUT * H * H + COUT * 4 * 4  # U8 spikes in, spikes out, pooled out
PEAK_FP32 = 157.3e12
PEAK_FP16 = 2.5e15
PEAK_INT8 = 5.0e15
PEAK_HBM = 8.0e12
# HBM bytes one full-batch whole-sample launch moves on this workload, from rocprofv3
# FETCH_SIZE (x2.0, the factor a 512 MiB device copy reads as) + WRITE_SIZE (x1.0):
# profiles/fused_layer8_c2_traffic.txt (int8) and profiles/fused_layer_c2_traffic.txt (fp16).
# Algorithmic: 90.3 MB input + 9.6 MB summaries/first-spike map + 814.0 MB spikes and pooled spikes
# (the kernel writes them) = 914.0 MB; the spike lines that get 1s are written twice (non-temporal
# zeros, then the 1s).
TRAFFIC_FUSED_LAYER = {"fp16": 979.2e6, "int8": 985.7e6}
# The region kernel per launch on the precision workload, and the whole batching-workload step
# (region kernel, writers, convwta), from the same counters: profiles/fused_region_traffic.txt.
TRAFFIC_REGION_C3 = 29.16e9
TRAFFIC_STEP_C4 = 79.7e9
WARM_S = 0.3  # seconds of untimed fused steps before the timed ones (at least --warmup steps)
METRIC = "images\u00b7timesteps/s, conv+fire+wta+STDP layer: FP16-batched vs FP32 plain, 1 GPU"


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
    return spk.pool(out, 3, 3), win, out


def fused_step(spk, x, conv):
    out, win, pooled = spk.layer_step(conv, x, TH, RADIUS, COUNT, 3, 3)
    return pooled, win, out


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def precision(spk, x, k2, dev, samples=64):
    """The fused path vs the FP64 check conv and vs the FP32 plain step."""
    import torch
    from spyker_amd import _lib
    xs = x[:samples].contiguous()
    n = xs.shape[0]
    kd = torch.from_numpy(k2).to(dev)
    p64 = torch.empty((n, T_STEPS, COUT, H, H), dtype=torch.float64, device=dev)
    spk.call("spk_conv", spk._ptr(xs), _lib.U8, spk._ptr(kd), _lib.F32, spk._ptr(p64), _lib.F64, n * T_STEPS, CIN,
             H, H, COUT, K, K, 1, 1, 1, 1, 1, 1, spk._stream(xs))
    conv = spk.Conv(CIN, COUT, K, pad=1, device=dev)
    conv.kernel = kd.clone()
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    p16 = torch.empty((n, T_STEPS, COUT, H, H), dtype=torch.float32, device=dev)
    out_f, win_f, _ = spk.layer_step(conv, xs, TH, RADIUS, COUNT, 3, 3, potentials=p16)
    conv_p = spk.Conv(CIN, COUT, K, pad=1, device=dev)
    conv_p.kernel = kd.clone()
    conv_p.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    _, win_p, out_p = plain_step(spk, xs, conv_p)
    torch.cuda.synchronize()
    flips = ((p16.double() > TH) != (p64 > TH)).double().mean().item()
    wf, wp = win_f.data[:, :, 2:].cpu(), win_p.data[:, :, 2:].cpu()
    cf, cp = win_f.counts.cpu(), win_p.counts.cpu()
    same = sum(int(cf[i] == cp[i] and bool((wf[i, :cf[i]] == wp[i, :cp[i]]).all())) for i in range(n))
    return {
        "samples": n,
        "threshold_flip_rate_vs_fp64": flips,
        "max_abs_potential_error_vs_fp64": (p16.double() - p64).abs().max().item(),
        "spike_mismatch_rate_vs_fp32_plain": (out_f != out_p).double().mean().item(),
        "samples_with_identical_winners_vs_fp32_plain": same,
        "max_abs_weight_diff_vs_fp32_plain_after_step": (conv.kernel - conv_p.kernel).abs().max().item(),
    }


def host_cpu():
    """The host CPU: model, physical cores, logical CPUs, and the CPUs this process may use."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for ln in f:
                k, _, v = ln.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "usable_cpus": usable}


def _omp_threads(n):
    """Sets the OpenMP thread count of the host libraries (libgomp) and torch."""
    import torch
    try:
        C.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass
    torch.set_num_threads(int(n))


def _cpu_step(R, x, k2, conv):
    """One layer step on the host; returns the seconds each stage took."""
    n = x.shape[0]
    t = {}
    t0 = time.perf_counter()
    pot = conv(x)
    t["conv"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    R.threshold(pot, TH, inplace=True)
    R.inhibit(pot, inplace=True)
    w, c = R.convwta(pot, RADIUS, COUNT)
    out = R.fire(pot)
    t["threshold+inhibit+convwta+fire"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    R.convstdp(x, k2, out, [(0.004, -0.003)], w, c, 1, (1, 1, 1, 1))
    t["stdp"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    R.pool(out, 3, 3)
    t["pool"] = time.perf_counter() - t0
    del pot, out
    return t


def cpu_baseline(x_host, k2, samples, one_core_samples=64, canonical_samples=64):
    """The reference's CPU path of the same layer step on the host cores.

    Every stage after the conv is the reference's own CPU code (cpu/fire.cpp,
    inhibit.cpp, wta.cpp, stdp.cpp, pool.cpp, compiled from its sources into
    oracle/_ref); the conv, which the reference runs through oneDNN (absent
    here), is torch's CPU conv2d, itself oneDNN, on the spikes cast to F32 as
    opers.cpp:410 casts them.  Timed on all usable host CPUs and on one, and
    with the oracle's canonical direct conv in place of oneDNN."""
    import oracle
    import torch
    orc = oracle.Oracle()
    ref = oracle.reference_available()
    R = oracle.Reference() if ref else orc
    cpu = host_cpu()
    threads = cpu["usable_cpus"]
    kt = torch.from_numpy(k2)

    def onednn(x):
        n = x.shape[0]
        xf = torch.from_numpy(x.reshape(n * T_STEPS, CIN, H, H)).float()
        with torch.no_grad():
            pot = torch.nn.functional.conv2d(xf, kt, padding=1)
        return np.ascontiguousarray(pot.numpy().reshape(n, T_STEPS, COUT, H, H))

    def canonical(x):
        n = x.shape[0]
        return orc.conv(x.reshape(n * T_STEPS, CIN, H, H), k2, 1, (1, 1, 1, 1),
                        threads=threads).reshape(n, T_STEPS, COUT, H, H)

    def run(x, conv, nthreads, reps=1):
        _omp_threads(nthreads)
        _cpu_step(R, x[:2], k2, conv)  # warm (pages, thread pools)
        best = None
        for _ in range(reps):
            t = _cpu_step(R, x, k2, conv)
            if best is None or sum(t.values()) < sum(best.values()):
                best = t
        return best

    x = np.ascontiguousarray(x_host[:samples])
    n = x.shape[0]
    t_all = run(x, onednn, threads, reps=2)
    x1 = np.ascontiguousarray(x_host[:one_core_samples])
    t_one = run(x1, onednn, 1)
    xc = np.ascontiguousarray(x_host[:canonical_samples])
    t_can = run(xc, canonical, threads)
    _omp_threads(threads)
    rate = lambda m, t: m * T_STEPS / sum(t.values())
    what = "the reference's own CPU stages (oracle/_ref)" if ref else "the oracle's C port of the stages"
    return {"value": rate(n, t_all), "unit": "images*timesteps/s", "cores": threads,
            "kind": "reference" if ref else "port",
            "sample": "%d images x %d steps of the same layer step: torch CPU conv2d (oneDNN, standing in for the "
                      "reference's oneDNN conv) + %s for threshold, inhibit, convwta, fire, STDP, pool; %d threads, "
                      "best of 2" % (n, T_STEPS, what, threads),
            "cpu": cpu, "stage_seconds": {k: round(v, 4) for k, v in t_all.items()},
            "one_core": {"value": rate(len(x1), t_one), "threads": 1, "images": len(x1),
                         "stage_seconds": {k: round(v, 4) for k, v in t_one.items()}},
            "with_canonical_conv": {"value": rate(len(xc), t_can), "threads": threads, "images": len(xc),
                                    "conv": "the oracle's direct loop (the FP32 parity order)",
                                    "stage_seconds": {k: round(v, 4) for k, v in t_can.items()}}}


SP_N, SP_BINS, SP_SIZE, SP_COUT, SP_K, SP_TH, SP_COUNT = 256, 20, 128, 64, 5, 8.0, 16


def sparse_inputs(dev, n=SP_N, seed=5):
    import torch
    from spyker_amd import synthetic
    ev = synthetic.events(n, SP_BINS, SP_SIZE, seed=seed, density=0.02)
    rng = np.random.Generator(np.random.PCG64(seed))
    k = np.clip(rng.normal(0.8, 0.05, (SP_COUT, 2, SP_K, SP_K)), 0, 1).astype(np.float32)
    return ev, torch.from_numpy(ev).to(dev), k


def sparse_step(spk, ev, conv):
    """events -> sparse conv -> fcwta -> sparse STDP.  The fcwta runs on the
    lists and returns the winners fcwta(out.dense()) picks, each with its
    (z, y, x) and spike step, which the sparse STDP takes directly."""
    s = spk.SparseTensor(ev)
    out = conv(s, SP_TH)
    win = spk.sparse.fcwta(out, 0, SP_COUNT)
    conv.stdp(s, win)
    return out, win


def sparse_dense_step(spk, ev, conv):
    """The same step with fcwta on the densified output (the reference's only route)."""
    s = spk.SparseTensor(ev)
    out = conv(s, SP_TH)
    n, hw = ev.shape[0], SP_SIZE * SP_SIZE
    win = spk.fcwta(out.dense().reshape(n, SP_BINS, -1), 0, SP_COUNT)
    first = spk.sparse.gather(out, "i32").reshape(n, -1)
    idx = win.data[:, :, 4]
    d = win.data.clone()
    d[:, :, 1], d[:, :, 2], d[:, :, 3], d[:, :, 4] = first.gather(1, idx).long(), idx // hw, \
        (idx % hw) // SP_SIZE, idx % SP_SIZE
    conv.stdp(s, spk.Winners(d, win.counts))
    return out, win


def sparse_bench(spk, dev, steps, warmup, cpu):
    import torch
    ev_host, ev, k = sparse_inputs(dev)
    conv = spk.Conv(2, SP_COUT, SP_K, pad=2, device=dev)
    conv.kernel = torch.from_numpy(k).to(dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    dt = timed(lambda: sparse_step(spk, ev, conv), steps, warmup)
    dtd = timed(lambda: sparse_dense_step(spk, ev, conv), 2, 1)
    out, _ = sparse_step(spk, ev, conv)
    s = spk.SparseTensor(ev)
    res = {"workload": "256 DVS-like samples, 2x128x128, 20 bins, 2% density; sparse conv 2->64 5x5 pad 2, "
                       "threshold 8, fcwta k16 (the winners of fcwta on the densified output, picked on the lists), "
                       "sparse STDP",
           "value": SP_N * SP_BINS / dt, "unit": "samples*bins/s", "ms_per_step": dt * 1e3, "steps": steps,
           "input_events_per_sample": s.numel / SP_N, "output_events_per_sample": out.numel / SP_N,
           "dtype": "f32", "ms_per_step_with_densified_fcwta": dtd * 1e3}
    if cpu:
        res["cpu_baseline"] = sparse_cpu_baseline(ev_host, k, 32)
    return res


def sparse_cpu_baseline(ev, k, samples):
    """The reference's own sparse CPU code (oracle/_ref) on a sample of the same workload."""
    import oracle
    if oracle.reference_available():
        X, kind, what = oracle.Reference(), "reference", "reference sparse code (oracle/_ref, OpenMP)"
    else:
        X, kind, what = oracle.Oracle(), "port", "oracle C port"
    ev = np.ascontiguousarray(ev[:samples])
    n = ev.shape[0]
    t0 = time.perf_counter()
    s = X.sparse_from_dense(ev)
    out = X.sparse_conv(s, k, SP_TH, (2, 2, 2, 2))
    w, c = X.fcwta(X.sparse_to_dense(out).reshape(n, SP_BINS, -1), 0, SP_COUNT)
    dt = time.perf_counter() - t0  # STDP on the CPU is left out: it is a small share there
    return {"value": n * SP_BINS / dt, "unit": "samples*bins/s", "cores": os.cpu_count() or 1, "kind": kind,
            "sample": "%d samples x %d bins: to-sparse, sparse conv, densify, fcwta (%s; STDP not timed)"
                      % (n, SP_BINS, what)}


def guarded(fn):
    """Runs one extra workload; a failure is reported in its place, never lost."""
    import torch
    try:
        return fn()
    except Exception as e:  # the headline line must still print
        return {"error": "%s: %s" % (type(e).__name__, e)}
    finally:
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


# Precision variant: 512 bar images 128x128, Gabor 4 orientations x {7x7 lambda 4, 11x11 lambda 8},
# gamma 0.5 -> 8 channels, rank code T=30, conv 8->64 7x7 pad 3, threshold 20, inhibit,
# convwta r5 k10, pool 2x2, STDP to the FP32 masters.
C3_N, C3_T, C3_S, C3_COUT, C3_K, C3_TH = 512, 30, 128, 64, 7, 20.0


def c3_inputs(spk, dev, n=C3_N, seed=3):
    import torch
    from spyker_amd import synthetic
    img = torch.from_numpy(synthetic.bars(n, C3_S, seed=seed)).to(dev).reshape(n, 1, C3_S, C3_S)
    banks = []
    for size, lam in ((3, 4.0), (5, 8.0)):
        filt = [spk.GaborFilter(0.56 * lam, th, 0.5, lam, 0.0) for th in (0.0, 45.0, 90.0, 135.0)]
        banks.append(spk.Gabor(size, filt, pad=size, device=dev)(img))
    g = torch.cat(banks, dim=1).contiguous()  # [n, 8, 128, 128]
    spk.threshold(g, 0.0)
    x = spk.code(g, C3_T)  # [n, 30, 8, 128, 128]
    rng = np.random.Generator(np.random.PCG64(seed))
    k = np.clip(rng.normal(0.8, 0.05, (C3_COUT, 8, C3_K, C3_K)), 0, 1).astype(np.float32)
    return x, k


def pick(plan, choice):
    """Fixes the plan's candidate to (rows a region, write mode), when given."""
    if choice:
        cands, _, _ = plan.candidates()
        rows = [c[0] for c in cands]
        if choice[0] not in rows:
            raise SystemExit("no plan candidate with %d rows (candidates: %s)" % (choice[0], rows))
        plan.select(rows.index(choice[0]), choice[1])


def c3_bench(spk, dev, steps, choice=None):
    import torch
    x, k = c3_inputs(spk, dev)
    n = x.shape[0]
    conv = spk.Conv(8, C3_COUT, C3_K, pad=3, device=dev)
    conv.kernel = torch.from_numpy(k).to(dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    plan = spk.LayerPlan(conv, [(C3_T, C3_S, C3_S)] * n, C3_TH, 5, 10, 2, 2, tune=True)
    packed = x.reshape(-1)
    bufs = plan.buffers()
    plan.run(packed, buffers=bufs)  # tunes the region height
    pick(plan, choice)
    spk.call("spk_kernel_timer", 1)
    dt = timed(lambda: plan.run(packed, buffers=bufs), steps, 1)
    total_ms, launches = C.c_double(0), C.c_int64(0)
    spk.call("spk_kernel_timer_read", C.byref(total_ms), C.byref(launches))
    spk.call("spk_kernel_timer", 0)
    kern_s = total_ms.value / 1e3 / max(launches.value, 1)
    flop = 2.0 * C3_COUT * C3_S * C3_S * 8 * C3_K * C3_K * n * C3_T
    res = {"workload": "512 bar images 128x128, Gabor 4x2 -> 8 ch, T=30, conv 8->64 7x7 pad 3, threshold 20, "
                       "inhibit, convwta r5 k10, pool 2x2, STDP",
           "value": n * C3_T / dt, "unit": "images*timesteps/s", "ms_per_step": dt * 1e3, "steps": steps,
           "dtype": "f16-mfma/f32-acc",
           "roofline": {"bound": "mfma", "achieved": flop / kern_s / 1e12, "peak": PEAK_FP16 / 1e12,
                        "unit": "TFLOP/s", "frac": flop / kern_s / PEAK_FP16, "traffic": TRAFFIC_REGION_C3,
                        "kernel": "fused_region", "kernel_ms": kern_s * 1e3}}
    cands, chosen, writes = plan.candidates()
    res["region_kernel"] = {"rows_per_region": cands[chosen][0], "writes_in_kernel": bool(writes)}
    # precision against the FP64 conv on a few samples
    m = 8
    xs = x[:m].contiguous()
    p64 = torch.empty((m, C3_T, C3_COUT, C3_S, C3_S), dtype=torch.float64, device=dev)
    kd = torch.from_numpy(k).to(dev)
    spk.call("spk_conv", spk._ptr(xs), 4, spk._ptr(kd), 9, spk._ptr(p64), 10, m * C3_T, 8, C3_S, C3_S, C3_COUT,
             C3_K, C3_K, 1, 1, 3, 3, 3, 3, spk._stream(xs))
    cp = spk.Conv(8, C3_COUT, C3_K, pad=3, device=dev)
    cp.kernel = kd.clone()
    cp.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    p16 = torch.empty((m, C3_T, C3_COUT, C3_S, C3_S), dtype=torch.float32, device=dev)
    out16, win16, _ = spk.layer_step(cp, xs, C3_TH, 5, 10, learn=True, potentials=p16)
    flips = ((p16.double() > C3_TH) != (p64 > C3_TH)).double().mean().item()
    perr = (p16.double() - p64).abs().max().item()
    del p64, p16
    # the FP32 plain path (the reference's semantics) on the same samples and weights
    cf = spk.Conv(8, C3_COUT, C3_K, pad=3, device=dev)
    cf.kernel = kd.clone()
    cf.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    pot = cf(xs)
    spk.threshold(pot, C3_TH)
    spk.inhibit(pot)
    winp = spk.convwta(pot, 5, 10)
    outp = spk.fire(pot)
    del pot
    cf.stdp(xs, winp, outp)
    wf, wp = win16.data[:, :, 2:].cpu(), winp.data[:, :, 2:].cpu()
    cf_, cp_ = win16.counts.cpu(), winp.counts.cpu()
    same = sum(int(cf_[i] == cp_[i] and bool((wf[i, :cf_[i]] == wp[i, :cp_[i]]).all())) for i in range(m))
    res["precision"] = {"samples": m, "threshold_flip_rate_vs_fp64": flips, "max_abs_potential_error_vs_fp64": perr,
                        "spike_mismatch_rate_vs_fp32_plain": (out16 != outp).double().mean().item(),
                        "samples_with_identical_winners_vs_fp32_plain": same,
                        "max_abs_weight_diff_vs_fp32_plain_after_step": (cp.kernel - cf.kernel).abs().max().item()}
    return res


# Batching variant: 4096 stroke images of uneven size S in {32..256} and T in 8..32, DoG 4 ch,
# conv 4->32 5x5 pad 2, threshold 12, inhibit, convwta r2 k4, pool 2x2, STDP; one plan,
# one captured graph per step.
C4_N, C4_SIZES, C4_COUT, C4_TH = 4096, (32, 48, 64, 96, 128, 192, 256), 32, 12.0


def c4_inputs(spk, dev, n=C4_N, seed=4):
    import torch
    from spyker_amd import synthetic
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice(C4_SIZES, n)
    times = rng.integers(8, 33, n)
    shapes = [(int(t), int(s), int(s)) for t, s in zip(times, sizes)]
    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    coded = [None] * n
    for s in C4_SIZES:
        idx = np.nonzero(sizes == s)[0]
        if len(idx) == 0:
            continue
        img = torch.from_numpy(synthetic.strokes(len(idx), int(s), seed=seed + int(s)).astype(np.float32) / 255)
        img = img.to(dev).reshape(len(idx), 1, int(s), int(s))
        f = torch.cat([d3(img), d7(img)], dim=1).contiguous()
        spk.threshold(f, 0.01)
        for t in np.unique(times[idx]):
            sel = idx[times[idx] == t]
            pos = np.searchsorted(idx, sel)
            x = spk.code(f[torch.from_numpy(pos).to(dev)].contiguous(), int(t))
            for j, i in enumerate(sel):
                coded[i] = x[j]
    rng2 = np.random.Generator(np.random.PCG64(seed))
    k = np.clip(rng2.normal(0.8, 0.05, (C4_COUT, 4, 5, 5)), 0, 1).astype(np.float32)
    return shapes, coded, k


def c4_bench(spk, dev, steps, choice=None):
    import torch
    shapes, coded, k = c4_inputs(spk, dev)
    conv = spk.Conv(4, C4_COUT, 5, pad=2, device=dev)
    conv.kernel = torch.from_numpy(k).to(dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    plan = spk.LayerPlan(conv, shapes, C4_TH, 2, 4, 2, 2, tune=True)
    packed = plan.pack(coded)
    del coded
    bufs = plan.buffers()
    plan.run(packed, buffers=bufs)  # tunes the region height, configures the kernels
    pick(plan, choice)
    dt = timed(lambda: plan.run(packed, buffers=bufs), steps, 1)
    # the same step captured once and replayed (the launches are few, so a
    # replay saves little; the graph executor's queue assignment can order a
    # chunk's kernel behind an earlier chunk's writers)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(packed, buffers=bufs)
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        plan.run(packed, buffers=bufs)
    dt_graph = timed(g.replay, steps, 1)
    spk.call("spk_kernel_timer", 1)
    plan.run(packed, buffers=bufs)
    torch.cuda.synchronize()
    total_ms, launches = C.c_double(0), C.c_int64(0)
    spk.call("spk_kernel_timer_read", C.byref(total_ms), C.byref(launches))
    spk.call("spk_kernel_timer", 0)
    kern_s = total_ms.value / 1e3 / max(launches.value, 1)
    its = sum(t for t, _, _ in shapes)
    flop = sum(2.0 * C4_COUT * h * w * 4 * 25 * t for t, h, w in shapes)
    # 145 FLOP per byte of U8 input and outputs, under the ~312 ridge: the step is bound by HBM
    step_bytes = plan.in_bytes + plan.out_bytes + plan.pool_bytes
    cands, chosen, writes = plan.candidates()
    return {"workload": "4096 stroke images, S in {32..256}, T in 8..32, DoG 4 ch, conv 4->32 5x5 pad 2, "
                        "threshold 12, inhibit, convwta r2 k4, pool 2x2, STDP; one plan",
            "value": its / dt, "unit": "images*timesteps/s", "ms_per_step": dt * 1e3, "steps": steps,
            "graph_replay_ms_per_step": dt_graph * 1e3,
            "images_timesteps": its, "dtype": "f16-mfma/f32-acc",
            "bytes_in_out_pooled": [plan.in_bytes, plan.out_bytes, plan.pool_bytes],
            "roofline": {"bound": "hbm", "achieved": step_bytes / dt / 1e9, "peak": PEAK_HBM / 1e9,
                         "unit": "GB/s", "frac": step_bytes / dt / PEAK_HBM, "traffic": TRAFFIC_STEP_C4,
                         "kernel": "the whole step (its HBM floor: the U8 input, spikes and pooled spikes)",
                         "floor_ms": step_bytes / PEAK_HBM * 1e3},
            "region_kernel": {"achieved_tflops": flop / kern_s / 1e12, "frac_of_fp16_mfma_peak": flop / kern_s / PEAK_FP16,
                              "kernel_ms": kern_s * 1e3, "rows_per_region": cands[chosen][0],
                              "writes_in_kernel": bool(writes)},
            "step_hbm_bytes_alg": step_bytes}


# Front ends before the layer: ZCA whitening of 50000 CIFAR-shaped samples (3x32x32 = 3072
# features, FP64 on the device), the LoG bank on the 1024 28x28 images (sizes 3, stds 1 and 2),
# the Poisson rate code of the 1024 DoG maps (4x28x28, T=15) and rate fire on layer-2 shaped
# potentials [1024, 15, 30*14*14].
ZCA_ROWS, ZCA_COLS, ZCA_EPS = 50000, 3072, 1e-2
PEAK_FP64 = 78.6e12


def _event_time(fn, reps):
    """Mean device time of fn on the current stream (HIP events), after one untimed call."""
    import torch
    fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / 1e3 / reps


def front_end_bench(spk, dev, cpu):
    import torch
    from spyker_amd import synthetic
    import oracle
    orc = oracle.Oracle() if cpu else None
    threads = os.cpu_count() or 1
    out = {}

    # ZCA: correlated synthetic data (rank-256 signal + noise), fit then whiten in place
    g = torch.Generator(device=dev).manual_seed(7)
    basis = torch.randn(256, ZCA_COLS, device=dev, generator=g) / 16
    x = (torch.randn(ZCA_ROWS, 256, device=dev, generator=g) @ basis
         + 0.1 * torch.randn(ZCA_ROWS, ZCA_COLS, device=dev, generator=g)).contiguous()
    z = spk.ZCA()
    t0 = time.perf_counter()
    z.fit(x, ZCA_EPS)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    y = x.clone()
    apply_s = _event_time(lambda: z(y), 3)
    flop = 2.0 * ZCA_ROWS * ZCA_COLS * ZCA_COLS
    out["zca"] = {"workload": "%d x %d F32 (CIFAR-shaped), eps %g, FP64 inside" % (ZCA_ROWS, ZCA_COLS, ZCA_EPS),
                  "fit_s": fit_s, "apply_ms": apply_s * 1e3, "value": ZCA_ROWS / apply_s, "unit": "samples/s",
                  "roofline": {"bound": "mfma", "achieved": flop / apply_s / 1e12, "peak": PEAK_FP64 / 1e12,
                               "unit": "TFLOP/s", "frac": flop / apply_s / PEAK_FP64,
                               "kernel": "apply: rocBLAS dgemm + centring/narrowing kernels"}}
    if cpu:
        xs = x[:5000].cpu().numpy()
        t0 = time.perf_counter()
        m, tr = orc.zca_fit(xs, ZCA_EPS)
        orc.zca_trans(xs, m, tr)
        dt = time.perf_counter() - t0
        out["zca"]["cpu_baseline"] = {"value": 5000 / dt, "unit": "samples/s (fit + apply)", "cores": threads,
                                      "kind": "port", "sample": "5000 x 3072: the oracle's numpy FP64 fit and "
                                      "apply (host BLAS/LAPACK threads)"}
    del x, y, basis, z
    torch.cuda.empty_cache()

    img = torch.from_numpy(synthetic.strokes(N, 28, seed=2).astype(np.float32) / 255).to(dev).reshape(N, 1, 28, 28)
    lg = spk.LoG(3, [1.0, 2.0], pad=3, device=dev)
    log_s = _event_time(lambda: lg(img), 10)
    lflop = 2.0 * N * 4 * 28 * 28 * 49
    lbytes = N * 28 * 28 * 4 * (1 + 4)  # the F32 input plane and the 4 output planes of an image
    out["log"] = {"workload": "1024 x 28x28, LoG size 3 (7x7), stds 1 and 2, pad 3 -> [1024, 4, 28, 28]",
                  "ms": log_s * 1e3, "value": N / log_s, "unit": "images/s",
                  "roofline": {"bound": "hbm", "achieved": lbytes / log_s / 1e9,
                               "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": lbytes / log_s / PEAK_HBM,
                               "flop_rate_tflops": lflop / log_s / 1e12}}
    if cpu:
        ih, kh = img.cpu().numpy(), lg.kernel.cpu().numpy()
        t0 = time.perf_counter()
        orc.log(ih, kh, (3, 3, 3, 3))
        dt = time.perf_counter() - t0
        out["log"]["cpu_baseline"] = {"value": N / dt, "unit": "images/s", "cores": threads, "kind": "port",
                                      "sample": "the same 1024 images (oracle C port, OpenMP)"}

    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    filt = torch.cat([d3(img), d7(img)], dim=1).contiguous()
    spk.threshold(filt, 0.01)
    ln = 4 * 28 * 28
    rate_s = _event_time(lambda: spk.code(filt, T_STEPS, code="rate"), 10)
    rbytes = N * ln * (4 + 2 + 2 + T_STEPS)  # F32 in, U16 times out and back in, U8 counts per step
    out["rate_code"] = {"workload": "1024 DoG maps 4x28x28, T=15, sorted order (the default), U8 counts",
                        "ms": rate_s * 1e3, "value": N * T_STEPS / rate_s, "unit": "images*timesteps/s",
                        "roofline": {"bound": "hbm", "achieved": rbytes / rate_s / 1e9, "peak": PEAK_HBM / 1e9,
                                     "unit": "GB/s", "frac": rbytes / rate_s / PEAK_HBM,
                                     "kernel": "rank-time sort + rate_code_kernel"}}
    if cpu:
        fh = filt.cpu().numpy().reshape(N, ln)[:256]
        tab = np.zeros((T_STEPS + 1, 1024), np.uint16)
        spk._lib.load().spk_poisson_table(tab.ctypes.data_as(C.c_void_p), C.c_int64(T_STEPS + 1), C.c_int64(1024))
        t0 = time.perf_counter()
        orc.rate_code(fh, T_STEPS, False, 12345, tab)
        dt = time.perf_counter() - t0
        out["rate_code"]["cpu_baseline"] = {"value": 256 * T_STEPS / dt, "unit": "images*timesteps/s", "cores": 1,
                                            "kind": "port",
                                            "sample": "256 of the maps (oracle C port, one thread)"}

    pot = torch.rand((N, T_STEPS, CIN * H * H), device=dev).cumsum(1).contiguous()
    fire_s = _event_time(lambda: spk.fire(pot, 0.5, code="rate"), 10)
    fbytes = pot.numel() * (4 + 1)
    out["rate_fire"] = {"workload": "[1024, 15, 30x14x14] F32 potentials -> U8 counts, threshold 0.5",
                        "ms": fire_s * 1e3, "value": N * T_STEPS / fire_s, "unit": "images*timesteps/s",
                        "roofline": {"bound": "hbm", "achieved": fbytes / fire_s / 1e9, "peak": PEAK_HBM / 1e9,
                                     "unit": "GB/s", "frac": fbytes / fire_s / PEAK_HBM, "kernel": "rate_fire_kernel"}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a step takes well under a millisecond: enough warm-up for the clocks and
    # the allocator to settle, and enough steps for a steady mean
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-samples", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--sparse-only", action="store_true", help="time only the sparse event-stream workload")
    ap.add_argument("--only", choices=["c3", "c4", "sparse", "front"], help="time only this extra workload")
    ap.add_argument("--no-extra", action="store_true", help="skip the precision and batching workloads")
    ap.add_argument("--arith", choices=["int8", "fp16"], default="int8",
                    help="the whole-sample kernel's arithmetic (int8 fixed point, the library default, or fp16)")
    ap.add_argument("--plan", type=int, nargs=2, metavar=("ROWS", "WRITES"),
                    help="with --only c3/c4: run the plan candidate of ROWS rows a region and write mode WRITES "
                         "instead of the tuner's pick (for counter runs, whose serialized dispatches skew the "
                         "tuner's timing)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (HBM counter calibration under rocprofv3)")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        src = torch.ones(a.copy_probe << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.copy_probe << 20}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the layer path runs on a single GPU "
              "(--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    import spyker_amd as spk

    dev = torch.device("cuda", 0)
    spk.set_fused_arith(a.arith)
    if a.sparse_only:
        a.only = "sparse"
    if a.only:
        fns = {"c3": lambda: c3_bench(spk, dev, a.steps, a.plan), "c4": lambda: c4_bench(spk, dev, a.steps, a.plan),
               "sparse": lambda: sparse_bench(spk, dev, a.steps, a.warmup, not a.no_cpu),
               "front": lambda: front_end_bench(spk, dev, not a.no_cpu)}
        print(json.dumps({a.only: fns[a.only]()}))
        return
    x, k2 = make_inputs(dev)

    def fused_run(arith):
        """Warm-up and timed fused steps in one arithmetic: (seconds a step, seconds a whole-sample launch,
        launches timed, the conv)."""
        spk.set_fused_arith(arith)
        conv = spk.Conv(CIN, COUT, K, pad=1, device=dev)
        conv.kernel = torch.from_numpy(k2).to(dev)
        conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
        # W warm-up steps, and at least WARM_S seconds of them: a step is half a
        # millisecond, and the GPU's clocks take a few hundred steps to settle
        t_end = time.perf_counter() + WARM_S
        i = 0
        while i < a.warmup or time.perf_counter() < t_end:
            fused_step(spk, x, conv)
            i += 1
            if i % 32 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        spk.call("spk_kernel_timer", 1)
        d = timed(lambda: fused_step(spk, x, conv), a.steps, 0)
        total, nl = C.c_double(0), C.c_int64(0)
        spk.call("spk_kernel_timer_read", C.byref(total), C.byref(nl))
        spk.call("spk_kernel_timer", 0)
        return d, total.value / 1e3 / max(nl.value, 1), nl.value, conv

    dt, kern_s, nlaunch, conv_f = fused_run(a.arith)
    launches = C.c_int64(nlaunch)
    achieved = FLOP_PER_IT * N * T_STEPS / kern_s
    if a.arith == "int8":
        # spike bytes times 16-bit fixed-point weights, split into a high and a low int8 byte:
        # two int8 MFMA products per multiply-add of the conv, I32 accumulation
        dtype, peak, kname = "u8*i16-fixed (int8 mfma, i32 acc)", PEAK_INT8, "fused_layer8"
        path = "fused int8-MFMA layer step (16-bit fixed-point weights), FP32 master weights"
    else:
        dtype, peak, kname = "f16-mfma/f32-acc", PEAK_FP16, "fused_layer"
        path = "fused FP16-MFMA layer step, FP32 master weights"

    line = {
        "metric": json.loads('"%s"' % METRIC),
        "value": N * T_STEPS / dt,
        "unit": "images*timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": "layer 2 of the two-layer stroke-image STDP network: batch 1024 of 28x28 images, "
                               "T=15, conv 30->250 3x3 pad 1 on 14x14 spikes, threshold 10, inhibit, convwta r1 k8, "
                               "STDP, pool 3x3 s3",
                   "path": path},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": TRAFFIC_FUSED_LAYER[a.arith], "traffic_unit": "bytes/launch",
                     "kernel": kname, "kernel_ms": kern_s * 1e3, "launches_timed": launches.value,
                     "flop_per_launch": FLOP_PER_IT * N * T_STEPS},
    }
    line["config"]["channels_per_wave"] = spk.fused_tuned_channels(conv_f, T_STEPS, H, H)
    # the other arithmetic of the same kernel, timed the same way on the same inputs
    other = "fp16" if a.arith == "int8" else "int8"
    dto, kso, _, _ = fused_run(other)
    spk.set_fused_arith(a.arith)
    line["other_arithmetic"] = {
        "arith": other, "value": N * T_STEPS / dto, "ms_per_step": dto * 1e3, "kernel_ms": kso * 1e3,
        "achieved_tflops": FLOP_PER_IT * N * T_STEPS / kso / 1e12,
        "frac_of_fp16_mfma_peak": FLOP_PER_IT * N * T_STEPS / kso / PEAK_FP16}
    if a.arith == "int8":
        line["roofline"]["note"] = ("achieved counts the conv's own multiply-adds; the int8 path issues two int8 "
                                    "products for each (high and low weight byte), so its ceiling for this conv "
                                    "is half the int8 peak, the FP16 peak")
        line["roofline"]["frac_of_two_product_ceiling"] = achieved / (PEAK_INT8 / 2)
    if not a.no_plain:
        conv_p = spk.Conv(CIN, COUT, K, pad=1, device=dev)
        conv_p.kernel = torch.from_numpy(k2).to(dev)
        conv_p.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
        dtp = timed(lambda: plain_step(spk, x, conv_p), a.plain_steps, 1)
        line["plain"] = {"value": N * T_STEPS / dtp, "ms_per_step": dtp * 1e3, "dtype": "f32",
                         "speedup_fused_over_plain": dtp / dt}
    line["precision"] = precision(spk, x, k2, dev)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(x.cpu().numpy(), k2, a.cpu_samples)
    del x
    torch.cuda.empty_cache()
    if not a.no_sparse:
        line["sparse"] = guarded(lambda: sparse_bench(spk, dev, 5, 1, not a.no_cpu))
    if not a.no_extra:
        line["precision_variant"] = guarded(lambda: c3_bench(spk, dev, 3))
        line["batching_variant"] = guarded(lambda: c4_bench(spk, dev, 3))
        line["front_end"] = guarded(lambda: front_end_bench(spk, dev, not a.no_cpu))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
