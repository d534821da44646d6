"""The fused layer step (int8 fixed-point and FP16 MFMA arithmetic).

Two bars:
  * structure: with weights exactly representable in FP16 and potentials
    that sum exactly in FP32 (multiples of 1/64), the FP16 path computes the
    same potentials as the FP32 plain path, so spikes, winners, pooled spikes
    and STDP-updated weights must equal the oracle's bit for bit;
  * precision: with N(0.8, 0.05) weights, raw potentials stay within
    K * 2^-11 * max|w| of FP64 and the threshold-flip rate stays <= 0.1 %.
"""

import contextlib

import numpy as np
import pytest
import torch

import spyker_amd as spk

pytestmark = pytest.mark.gpu

ARITH = ["int8", "fp16"]


@contextlib.contextmanager
def arith(mode):
    """The whole-sample kernel's arithmetic for the duration of a block (int8 is the default)."""
    spk.set_fused_arith(mode)
    try:
        yield
    finally:
        spk.set_fused_arith("int8")


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


def oracle_step(orc, x, k, th, radius, count, pad, pool, config=(0.004, -0.003)):
    n, t, cin, h, w = x.shape
    cout, _, kh, kw = k.shape
    pt, pl, pb, pr = pad
    ho, wo = h + pt + pb - kh + 1, w + pl + pr - kw + 1
    pot = orc.conv(x.reshape(n * t, cin, h, w), k, 1, pad).reshape(n, t, cout, ho, wo)
    pot = orc.inhibit(orc.threshold(pot, th))
    win, cnt = orc.convwta(pot, radius, count)
    out = orc.fire(pot)
    k2 = orc.convstdp(x, k, out, [config], win, cnt, 1, pad)
    pooled = orc.pool(out, pool[0], pool[1]) if pool else None
    return out, win, cnt, k2, pooled


def spikes(rng, n, t, c, h, w, p=0.15, cumulative=True):
    s = (rng.random((n, t, c, h, w)) < p / t * 3).astype(np.uint8)
    return np.maximum.accumulate(s, axis=1) if cumulative else s


def check_winners(wn, win, cnt):
    d, c = H(wn.data), H(wn.counts)
    assert np.array_equal(c, cnt)
    for i in range(len(c)):
        for k in range(c[i]):
            assert tuple(d[i, k, 2:]) == (win[i, k]["z"], win[i, k]["y"], win[i, k]["x"])


@pytest.mark.parametrize("case", [
    # n, t, cin, h, w, cout, k, pad, th, radius, count, pool, cumulative input
    (6, 15, 30, 14, 14, 250, 3, (1, 1, 1, 1), 10.0, 1, 8, (3, 3), True),
    (5, 15, 4, 28, 28, 30, 5, (2, 2, 2, 2), 15.0, 3, 5, (2, 2), True),
    (3, 8, 8, 20, 17, 64, 7, (3, 3, 3, 3), 20.0, 5, 10, (2, 2), True),
    (4, 12, 3, 11, 9, 40, 3, (0, 1, 2, 0), 4.0, (1, 2), 4, None, True),
    # isolated spikes: potentials rise and fall, fire masks have gaps
    (4, 10, 16, 12, 12, 48, 3, (1, 1, 1, 1), 6.0, 1, 6, (2, 2), False),
    # a plane of 1296 pixels: convwta by the block-per-sample kernel (the
    # smaller planes above take the wave-per-sample one)
    (2, 8, 4, 36, 36, 32, 3, (1, 1, 1, 1), 5.0, 2, 6, (2, 2), True),
    # the longest step counts the whole-sample kernel takes: 32-bit fire masks
    # at full width, with gaps (isolated spikes) and cumulative
    (3, 32, 8, 12, 12, 64, 3, (1, 1, 1, 1), 10.0, 1, 6, (2, 2), True),
    (3, 32, 8, 12, 12, 64, 3, (1, 1, 1, 1), 6.0, 1, 6, (2, 2), False),
    (3, 24, 16, 10, 10, 96, 3, (1, 1, 1, 1), 10.0, 2, 5, (2, 2), True),
])
@pytest.mark.parametrize("mode", ARITH)
def test_fused_exact_on_dyadic_weights(dev, orc, case, mode):
    n, t, cin, h, w, cout, k, pad, th, radius, count, pool, cumulative = case
    rng = np.random.default_rng(11)
    x = spikes(rng, n, t, cin, h, w, 0.5 if cumulative else 1.0, cumulative)
    kern = (rng.integers(40, 64, (cout, cin, k, k)) / 64).astype(np.float32)
    # scale the threshold to the potentials of this input
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64 if th else th
    conv = spk.Conv(cin, cout, k, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    with arith(mode):
        out, wn, pooled = spk.layer_step(conv, T(x, dev), th, radius, count, *(pool if pool else (None, None)))
    o_out, o_win, o_cnt, o_k, o_pool = oracle_step(orc, x, kern, th, radius, count, pad, pool)
    assert o_cnt.sum() > 0
    assert np.array_equal(H(out), o_out)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)
    if pool:
        assert np.array_equal(H(pooled), o_pool)


@pytest.mark.parametrize("case", [
    # n, t, cin, h, w, cout, k, pad, weight numerators lo..hi (of 64), byte density
    (4, 8, 4, 12, 12, 32, 3, (1, 1, 1, 1), (-64, 64), 0.3),
    # K = 270 with weights near 1: 255 * K * w * 2^15 sits just under 2^31,
    # the edge of the int8 kernel's range for its scale
    (3, 10, 30, 10, 10, 64, 3, (1, 1, 1, 1), (56, 64), 0.5),
    (3, 6, 8, 9, 11, 48, 5, (2, 2, 2, 2), (-20, 64), 0.2),
])
@pytest.mark.parametrize("mode", ARITH)
def test_fused_any_input_bytes(dev, orc, case, mode):
    """U8 inputs are not limited to spikes: conv sees every byte value (the
    reference casts them to float, opers.cpp:408-423), here 0..255 at random
    steps, with weights of both signs; on dyadic weights every sum is exact,
    so the fused step must equal the oracle bit for bit."""
    n, t, cin, h, w, cout, k, pad, (lo, hi), dens = case
    rng = np.random.default_rng(17)
    x = (rng.integers(1, 256, (n, t, cin, h, w)) * (rng.random((n, t, cin, h, w)) < dens)).astype(np.uint8)
    x[0, 0, 0, 0, :3] = 255
    kern = (rng.integers(lo, hi, (cout, cin, k, k)) / 64).astype(np.float32)
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    conv = spk.Conv(cin, cout, k, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    with arith(mode):
        out, wn, pooled = spk.layer_step(conv, T(x, dev), th, 1, 5, 2, 2)
    o_out, o_win, o_cnt, o_k, o_pool = oracle_step(orc, x, kern, th, 1, 5, pad, (2, 2))
    assert o_cnt.sum() > 0
    assert np.array_equal(H(out), o_out)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)
    assert np.array_equal(H(pooled), o_pool)


@pytest.mark.parametrize("config", [
    # positive, negative, stabilize, lower, upper
    (0.3, -0.2, True, 0.0, 1.0),      # lower +0: min/max then +0
    (0.3, -0.2, False, -0.5, 2.0),    # nonzero bounds: min/max
    (0.3, -0.25, True, 0.7, 0.8),     # nonzero bounds, clamps hit
    (0.004, -0.003, True, -1.0, 0.0), # upper 0: the reference's selects
    (0.2, -0.2, True, -0.0, 1.0),     # lower -0: the reference's selects
])
def test_fused_stdp_bounds(dev, orc, config):
    """The fused STDP chains against the oracle under every clamp form."""
    n, t, cin, h, w, cout, k, pad = 6, 15, 30, 14, 14, 250, 3, (1, 1, 1, 1)
    rng = np.random.default_rng(17)
    x = spikes(rng, n, t, cin, h, w, 0.5)
    kern = (rng.integers(40, 64, (cout, cin, k, k)) / 64).astype(np.float32)
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    conv = spk.Conv(cin, cout, k, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(*config)]
    out, wn, _ = spk.layer_step(conv, T(x, dev), th, 1, 8)
    o_out, o_win, o_cnt, o_k, _ = oracle_step(orc, x, kern, th, 1, 8, pad, None, config)
    assert o_cnt.sum() > 0
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel).view(np.uint32), o_k.view(np.uint32))


def test_fused_stdp_negative_zero_weights(dev, orc):
    """-0 weights under a +0 lower bound: the reference's clamp returns +0, so
    the F32 clamp (after the rounding) must order -0 below +0."""
    n, t, cin, h, w, cout, k, pad = 6, 15, 30, 14, 14, 250, 3, (1, 1, 1, 1)
    rng = np.random.default_rng(18)
    x = spikes(rng, n, t, cin, h, w, 0.5)
    kern = (rng.integers(40, 64, (cout, cin, k, k)) / 64).astype(np.float32)
    kern[:, :3] = -0.0
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    for config in [(0.3, -0.2, True, 0.0, 1.0), (0.3, -0.2, False, 0.0, 1.0)]:
        conv = spk.Conv(cin, cout, k, pad=pad, device=dev)
        conv.kernel = T(kern, dev)
        conv.stdpconfig = [spk.STDPConfig(*config)]
        _, wn, _ = spk.layer_step(conv, T(x, dev), th, 1, 8)
        _, o_win, o_cnt, o_k, _ = oracle_step(orc, x, kern, th, 1, 8, pad, None, config)
        assert o_cnt.sum() > 0
        check_winners(wn, o_win, o_cnt)
        got = H(conv.kernel)
        assert np.array_equal(got.view(np.uint32), o_k.view(np.uint32)), config
        # some -0 weights were updated to +0 (the case under test)
        assert ((o_k[:, :3] == 0) & ~np.signbit(o_k[:, :3])).any()


@pytest.mark.parametrize("mode", ARITH)
def test_fused_precision_vs_fp64(dev, orc, mode):
    n, t, cin, h, w, cout = 8, 15, 30, 14, 14, 250
    rng = np.random.default_rng(12)
    x = spikes(rng, n, t, cin, h, w, 0.5)
    kern = np.clip(rng.normal(0.8, 0.05, (cout, cin, 3, 3)), 0, 1).astype(np.float32)
    pot64 = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, (1, 1, 1, 1), f64=True).reshape(n, t, cout, h, w)
    th = float(np.quantile(pot64, 0.7))
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    pots = torch.empty((n, t, cout, h, w), dtype=torch.float32, device=dev)
    with arith(mode):
        spk.layer_step(conv, T(x, dev), th, 1, 8, learn=False, potentials=pots)
    p16 = H(pots).astype(np.float64)
    K = cin * 9
    assert np.abs(p16 - pot64).max() <= K * 2.0 ** -11 * kern.max()
    flips = np.mean((p16 > th) != (pot64 > th))
    assert flips <= 1e-3, flips


def test_fused_inference_only_leaves_weights(dev, orc):
    rng = np.random.default_rng(13)
    x = spikes(rng, 3, 10, 4, 16, 16, 0.5)
    kern = np.clip(rng.normal(0.8, 0.05, (20, 4, 5, 5)), 0, 1).astype(np.float32)
    conv = spk.Conv(4, 20, 5, pad=2, device=dev)
    conv.kernel = T(kern, dev)
    out, wn, _ = spk.layer_step(conv, T(x, dev), 5.0, 2, 3, learn=False)
    assert np.array_equal(H(conv.kernel), kern)
    assert H(out).sum() > 0


def test_fused_graph_replay(dev):
    """Replays of a captured layer step (the layer stack of a training loop
    in a hipGraph) give the eager steps' spikes, pooled spikes, winners and
    weights, step after step, the weights changing in place each step.

    The capture runs on the stream that was warmed, so the stream's workspace
    already has its size and the capture records no allocation; the workspace
    a capture used is kept for the life of the process."""
    rng = np.random.default_rng(41)
    n, t, cin, h, w, cout = 12, 15, 30, 14, 14, 250
    x = T(spikes(rng, n, t, cin, h, w), dev)
    kern = np.clip(rng.normal(0.8, 0.05, (cout, cin, 3, 3)), 0, 1).astype(np.float32)
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    eager = []
    for _ in range(3):
        out, wn, pooled = spk.layer_step(conv, x, 10.0, 1, 8, 3, 3)
        eager.append((out.clone(), pooled.clone(), wn.data.clone(), wn.counts.clone(), conv.kernel.clone()))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        spk.layer_step(conv, x, 10.0, 1, 8, 3, 3)  # warm on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    conv.kernel.copy_(T(kern, dev))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        out, wn, pooled = spk.layer_step(conv, x, 10.0, 1, 8, 3, 3)
    conv.kernel.copy_(T(kern, dev))
    torch.cuda.synchronize()
    for e in eager:
        g.replay()
        torch.cuda.synchronize()
        assert e[3].sum() > 0
        assert torch.equal(out, e[0]) and torch.equal(pooled, e[1])
        assert torch.equal(wn.data, e[2]) and torch.equal(wn.counts, e[3])
        assert torch.equal(conv.kernel, e[4])


@pytest.mark.parametrize("mode", ARITH)
@pytest.mark.parametrize("shape", [
    # n, t, cin, h, w, cout
    (6, 15, 30, 14, 14, 250),
    (5, 9, 4, 13, 11, 32),
    (4, 24, 12, 10, 9, 100),
])
def test_channels_a_wave_give_the_same_bytes(dev, orc, mode, shape):
    """32 and 64 output channels a wave (the tuner's two tile shapes) give the
    same spikes, pooled spikes, winners and weights, equal to the oracle's."""
    n, t, cin, h, w, cout = shape
    rng = np.random.default_rng(81)
    x = spikes(rng, n, t, cin, h, w, 0.35)
    kern = (rng.integers(40, 64, (cout, cin, 3, 3)) / 64).astype(np.float32)
    th = float(np.floor(0.3 * cin * 9 * 0.5))
    got = []
    with arith(mode):
        for ch in (32, 64):
            spk.set_fused_channels(ch)
            try:
                conv = spk.Conv(cin, cout, 3, pad=1, device=dev)
                conv.kernel = T(kern, dev)
                conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
                out, wn, pooled = spk.layer_step(conv, T(x, dev), th, 1, 5, 2, 2)
                got.append((H(out), H(pooled), H(wn.data), H(wn.counts), H(conv.kernel)))
            finally:
                spk.set_fused_channels(0)
    for a, b in zip(*got):
        assert np.array_equal(a, b)
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, (1, 1, 1, 1)).reshape(n, t, cout, h, w)
    pot = orc.inhibit(orc.threshold(pot, th))
    ow, oc = orc.convwta(pot, 1, 5)
    assert oc.sum() > 0 and np.array_equal(got[0][3], oc)
    assert np.array_equal(got[0][0], orc.fire(pot))


def test_tuner_keeps_a_choice_per_shape(dev):
    rng = np.random.default_rng(82)
    n, t, cin, h, w, cout = 40, 11, 8, 12, 12, 96
    x = T(spikes(rng, n, t, cin, h, w, 0.3), dev)
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev)
    conv.kernel = T((rng.integers(30, 64, (cout, cin, 3, 3)) / 64).astype(np.float32), dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    spk.set_fused_tuning(True)
    assert spk.fused_tuned_channels(conv, t, h, w) == 0
    first = spk.layer_step(conv, x, 20.0, 1, 4, learn=False)
    choice = spk.fused_tuned_channels(conv, t, h, w)
    assert choice in (32, 64)
    again = spk.layer_step(conv, x, 20.0, 1, 4, learn=False)
    assert torch.equal(first[0], again[0]) and torch.equal(first[1].data, again[1].data)
    assert spk.fused_tuned_channels(conv, t, h, w) == choice


def test_two_layer_stack_graph_replay(dev):
    """The two-layer stroke-image network's training step for layer 2 (layer
    1 forward: conv 4->30 5x5 pad 2, fire 15, pool 2; then the layer-2 step:
    conv 30->250 3x3 pad 1, threshold 10, inhibit, convwta r1 k8, fire,
    STDP, pool 3x3) captured as one hipGraph: every replay gives the eager
    steps' layer-1 spikes, spikes, pooled spikes, winners and weights."""
    from spyker_amd import synthetic
    n = 12
    img = torch.from_numpy(synthetic.strokes(n, 28, seed=43).astype(np.float32) / 255).to(dev).reshape(n, 1, 28, 28)
    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    filt = torch.cat([d3(img), d7(img)], dim=1).contiguous()
    spk.threshold(filt, 0.01)
    coded = spk.code(filt, 15)  # [n, 15, 4, 28, 28]
    rng = np.random.default_rng(44)
    k1 = T(np.clip(rng.normal(0.8, 0.05, (30, 4, 5, 5)), 0, 1).astype(np.float32), dev)
    k2 = np.clip(rng.normal(0.8, 0.05, (250, 30, 3, 3)), 0, 1).astype(np.float32)
    conv2 = spk.Conv(30, 250, 3, pad=1, device=dev)
    conv2.kernel = T(k2, dev)
    conv2.stdpconfig = [spk.STDPConfig(0.004, -0.003)]

    def stack():
        l1 = spk.pool(spk.fire(spk.conv(coded, k1, 1, 2), 15.0), 2)
        out, wn, pooled = spk.layer_step(conv2, l1, 10.0, 1, 8, 3, 3)
        return l1, out, wn, pooled

    eager = []
    for _ in range(3):
        l1, out, wn, pooled = stack()
        eager.append([v.clone() for v in (l1, out, pooled, wn.data, wn.counts, conv2.kernel)])
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        stack()  # warm the capture stream's workspaces
    torch.cuda.current_stream().wait_stream(side)
    conv2.kernel.copy_(T(k2, dev))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        l1, out, wn, pooled = stack()
    conv2.kernel.copy_(T(k2, dev))
    torch.cuda.synchronize()
    for e in eager:
        g.replay()
        torch.cuda.synchronize()
        assert e[4].sum() > 0
        for got, want in zip((l1, out, pooled, wn.data, wn.counts, conv2.kernel), e):
            assert torch.equal(got, want)


@pytest.mark.parametrize("case", [
    # n, t, cin, h, w, cout, k, stride, pad, th: each outside the fused step's domain
    (3, 10, 4, 20, 20, 16, 3, 2, (1, 1, 1, 1), 3.0),     # stride 2
    (2, 40, 4, 12, 12, 16, 3, 1, (1, 1, 1, 1), 3.0),     # 40 time steps
    (2, 8, 4, 10, 10, 300, 3, 1, (1, 1, 1, 1), 3.0),     # 300 output channels
    (2, 6, 64, 8, 8, 32, 3, 1, (1, 1, 1, 1), 20.0),      # Cin*Kh*Kw = 576
    (2, 8, 4, 12, 12, 16, 3, 1, (1, 1, 1, 1), -0.5),     # a negative threshold
])
def test_layer_step_falls_back_to_the_plain_path(dev, orc, case):
    """Outside the fused step's domain, layer_step runs the FP32 plain path,
    which matches the oracle's step bit for bit."""
    n, t, cin, h, w, cout, k, stride, pad, th = case
    rng = np.random.default_rng(19)
    x = spikes(rng, n, t, cin, h, w, 0.5)
    kern = np.clip(rng.normal(0.8, 0.05, (cout, cin, k, k)), 0, 1).astype(np.float32)
    conv = spk.Conv(cin, cout, k, stride=stride, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    ho, wo = (h + pad[0] + pad[2] - k) // stride + 1, (w + pad[1] + pad[3] - k) // stride + 1
    pots = torch.empty((n, t, cout, ho, wo), dtype=torch.float32, device=dev)
    out, wn, pooled = spk.layer_step(conv, T(x, dev), th, 1, 4, 2, 2, potentials=pots)
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, stride, pad).reshape(n, t, cout, ho, wo)
    assert np.array_equal(H(pots), pot)
    pot = orc.inhibit(orc.threshold(pot, th))
    o_win, o_cnt = orc.convwta(pot, 1, 4)
    o_out = orc.fire(pot)
    o_k = orc.convstdp(x, kern, o_out, [(0.004, -0.003)], o_win, o_cnt, stride, pad)
    assert o_cnt.sum() > 0
    assert np.array_equal(H(out), o_out)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)
    assert np.array_equal(H(pooled), orc.pool(o_out, 2, 2))
