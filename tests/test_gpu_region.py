"""The fused layer step over output regions (large samples, uneven batches).

Bars:
  * against the whole-sample fused kernel on a shape both run: identical
    spikes, pooled spikes, winners and STDP-updated weights, bit for bit
    (the same FP16 operands through the same MFMA sequence);
  * on weights exactly representable in FP16 whose potentials sum exactly in
    FP32, equal to the oracle's step bit for bit -- for uniform batches of
    large samples and for batches of uneven instances, where the oracle runs
    every instance's forward pass on the original weights and then the STDP
    of instance 0, 1, ... in order;
  * precision with N(0.8, 0.05) weights: potentials within K * 2^-11 * max|w|
    of FP64 and a threshold-flip rate <= 0.1 %;
  * a captured run (hipGraph through torch.cuda.graph) replays to the same
    results as eager runs.
"""

import numpy as np
import pytest
import torch

import spyker_amd as spk

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


def spikes(rng, shape, p=0.5, cumulative=True):
    s = (rng.random(shape) < p / shape[-4] * 3).astype(np.uint8)
    return np.maximum.accumulate(s, axis=-4) if cumulative else s


def dyadic(rng, shape):
    return (rng.integers(40, 64, shape) / 64).astype(np.float32)


def check_winners(wn, win, cnt, i=None):
    d, c = H(wn.data), H(wn.counts)
    rows = range(len(c)) if i is None else [i]
    for r in rows:
        g = 0 if i is not None else r
        assert c[r] == cnt[g], (r, c[r], cnt[g])
        for k in range(c[r]):
            assert tuple(d[r, k, 2:]) == (win[g, k]["z"], win[g, k]["y"], win[g, k]["x"]), (r, k)


def oracle_forward(orc, x, k, th, radius, count, pad, pool):
    t, cin, h, w = x.shape[-4:]
    n = x.shape[0]
    cout, _, kh, kw = k.shape
    pt, pl, pb, pr = pad
    ho, wo = h + pt + pb - kh + 1, w + pl + pr - kw + 1
    pot = orc.conv(x.reshape(n * t, cin, h, w), k, 1, pad).reshape(n, t, cout, ho, wo)
    pot = orc.inhibit(orc.threshold(pot, th))
    win, cnt = orc.convwta(pot, radius, count)
    out = orc.fire(pot)
    pooled = orc.pool(out, pool[0], pool[1]) if pool else None
    return out, win, cnt, pooled


def make_conv(dev, kern, pad):
    cout, cin, kh, kw = kern.shape
    conv = spk.Conv(cin, cout, (kh, kw), pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    return conv


@pytest.mark.parametrize("tune", [False, True])
def test_region_equals_whole_sample_kernel(dev, tune):
    """Both kernels in FP16 arithmetic sum the same operands in the same MFMA order."""
    n, t, cin, h, w, cout = 8, 15, 30, 14, 14, 250
    rng = np.random.default_rng(31)
    x = spikes(rng, (n, t, cin, h, w))
    kern = np.clip(rng.normal(0.8, 0.05, (cout, cin, 3, 3)), 0, 1).astype(np.float32)
    c1, c2 = make_conv(dev, kern, 1), make_conv(dev, kern, 1)
    spk.set_fused_arith("fp16")
    try:
        out1, wn1, pool1 = spk.layer_step(c1, T(x, dev), 10.0, 1, 8, 3, 3)
    finally:
        spk.set_fused_arith("int8")
    plan = spk.LayerPlan(c2, [(t, h, w)] * n, 10.0, 1, 8, 3, 3, tune=tune)
    out2, pool2, wn2 = plan.run(T(x, dev).reshape(-1))
    assert H(wn1.counts).sum() > 0
    assert torch.equal(out1.reshape(-1), out2) and torch.equal(pool1.reshape(-1), pool2)
    assert torch.equal(wn1.data, wn2.data) and torch.equal(wn1.counts, wn2.counts)
    assert torch.equal(c1.kernel, c2.kernel)


@pytest.mark.parametrize("case", [
    # n, t, cin, h, w, cout, k, pad, radius, count, pool, cumulative input
    (3, 10, 8, 40, 44, 64, 7, (3, 3, 3, 3), 5, 10, (2, 2), True),    # dword staging, 2 channel groups
    (2, 9, 8, 38, 42, 64, 7, (3, 3, 3, 3), 5, 10, (2, 2), True),     # w % 4 != 0: byte staging
    (3, 12, 4, 36, 40, 32, 5, (2, 1, 0, 3), 2, 4, (2, 2), True),     # one channel group, asymmetric pad
    (2, 7, 3, 33, 20, 100, 3, (1, 1, 1, 1), (1, 2), 6, None, True),  # 4 channel groups, ragged regions
    # isolated spikes: potentials rise and fall, so a pixel's fired steps
    # need not run to the last step and the tile end's gathered pass runs
    (3, 10, 8, 40, 44, 64, 7, (3, 3, 3, 3), 5, 10, (2, 2), False),
    (3, 12, 4, 36, 40, 32, 5, (2, 1, 0, 3), 2, 4, (2, 2), False),
])
def test_region_exact_on_dyadic_weights(dev, orc, case):
    n, t, cin, h, w, cout, k, pad, radius, count, pool, cumulative = case
    rng = np.random.default_rng(32)
    x = spikes(rng, (n, t, cin, h, w), 0.5 if cumulative else 1.5, cumulative)
    kern = dyadic(rng, (cout, cin, k, k))
    pot = orc.conv(x[:1].reshape(t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    conv = make_conv(dev, kern, pad)
    plan = spk.LayerPlan(conv, [(t, h, w)] * n, th, radius, count, *(pool if pool else (None, None)))
    out, pooled, wn = plan.run(T(x, dev).reshape(-1))
    o_out, o_win, o_cnt, o_pool = oracle_forward(orc, x, kern, th, radius, count, pad, pool)
    o_k = orc.convstdp(x, kern, o_out, [(0.004, -0.003)], o_win, o_cnt, 1, pad)
    assert o_cnt.sum() > 0
    # pixels whose fired steps (any channel) do not run to the last step take
    # the tile end's gathered pass; the isolated-spike cases must have some
    full = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad).reshape(n, t, cout, -1)
    fired = (full > th).any(axis=2)  # [n, t, pixels]
    suffix = np.flip(np.logical_and.accumulate(np.flip(fired, 1), axis=1), 1)
    assert (fired != suffix).any() != cumulative
    assert np.array_equal(H(out).reshape(o_out.shape), o_out)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)
    if pool:
        assert np.array_equal(H(pooled).reshape(o_pool.shape), o_pool)


def test_layer_step_takes_the_region_path_for_large_samples(dev, orc):
    """A sample too large for the LDS goes through spk_stdp_layer_fused unchanged."""
    n, t, cin, h, w, cout = 2, 12, 8, 48, 48, 64
    rng = np.random.default_rng(33)
    x = spikes(rng, (n, t, cin, h, w))
    kern = dyadic(rng, (cout, cin, 7, 7))
    pad = (3, 3, 3, 3)
    pot = orc.conv(x[:1].reshape(t, cin, h, w), kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    conv = make_conv(dev, kern, pad)
    out, wn, pooled = spk.layer_step(conv, T(x, dev), th, 5, 10, 2, 2)
    o_out, o_win, o_cnt, o_pool = oracle_forward(orc, x, kern, th, 5, 10, pad, (2, 2))
    o_k = orc.convstdp(x, kern, o_out, [(0.004, -0.003)], o_win, o_cnt, 1, pad)
    assert np.array_equal(H(out), o_out) and np.array_equal(H(pooled), o_pool)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)


@pytest.mark.parametrize("tune", [False, True])
def test_region_uneven_batch(dev, orc, tune):
    rng = np.random.default_rng(34)
    sizes = [16, 20, 24, 36, 40, 18, 32, 22]
    shapes = [(int(rng.integers(3, 10)), s, s) for s in sizes]
    cin, cout, k, pad, radius, count, pool = 4, 32, 5, (2, 2, 2, 2), 2, 4, (2, 2)
    xs = [spikes(rng, (t, cin, h, w)) for t, h, w in shapes]
    kern = dyadic(rng, (cout, cin, k, k))
    pot = orc.conv(xs[0], kern, 1, pad)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    conv = make_conv(dev, kern, pad)
    plan = spk.LayerPlan(conv, shapes, th, radius, count, *pool, tune=tune)
    out, pooled, wn = plan.run(plan.pack([T(x, dev) for x in xs]))
    o_k = kern
    for i, x in enumerate(xs):
        o_out, o_win, o_cnt, o_pool = oracle_forward(orc, x[None], kern, th, radius, count, pad, pool)
        assert np.array_equal(H(plan.spikes(out, i)), o_out[0]), i
        assert np.array_equal(H(plan.pooled(pooled, i)), o_pool[0]), i
        check_winners(wn, o_win, o_cnt, i)
        # STDP in the batch order: instance by instance, each on the kernel so far
        o_k = orc.convstdp(x[None], o_k, o_out, [(0.004, -0.003)], o_win, o_cnt, 1, pad)
    assert H(wn.counts).sum() > 0
    assert np.array_equal(H(conv.kernel), o_k)


def test_region_convwta_refills_candidates(dev, orc):
    """convwta where one channel fires first nearly everywhere: after that
    channel's winner, the later winners sit in threads whose listed
    candidates (their 8 best pixels, pixel p on thread p % 1024) all belong to
    that channel, so those threads find them only by refilling the list."""
    def sample(seed, t, h, w):
        rng = np.random.default_rng(seed)
        x = np.zeros((t, 4, h, w), np.uint8)
        x[:, 0] = rng.random((t, h, w)) < 0.35
        for _ in range(6):  # patches without channel 0 that start firing at step 3
            y0, x0, c = rng.integers(2, h - 14), rng.integers(2, w - 14), rng.integers(1, 4)
            x[:, 0, y0 - 2:y0 + 14, x0 - 2:x0 + 14] = 0
            x[3:, c, y0:y0 + 12, x0:x0 + 12] = rng.random((t - 3, 12, 12)) < 0.6
        return np.maximum.accumulate(x, axis=0)

    cout, pad, radius, count, pool = 32, (2, 2, 2, 2), 2, 6, (2, 2)
    kern = np.full((cout, 4, 5, 5), 8 / 64, np.float32)
    kern[0, 0] = 48 / 64
    for c in range(1, cout):
        kern[c, 1 + c % 3] = (40 + c) / 64
    shapes = [(8, 128, 128), (6, 100, 120)]
    xs = [sample(50 + i, *s) for i, s in enumerate(shapes)]
    th = float(np.quantile(orc.conv(xs[0], kern, 1, pad), 0.8)) + 0.5 / 64
    conv = make_conv(dev, kern, pad)
    plan = spk.LayerPlan(conv, shapes, th, radius, count, *pool, tune=False)
    out, pooled, wn = plan.run(plan.pack([T(x, dev) for x in xs]))
    o_k = kern
    for i, x in enumerate(xs):
        o_out, o_win, o_cnt, o_pool = oracle_forward(orc, x[None], kern, th, radius, count, pad, pool)
        # the case: winner 0 in channel 0 at step 0, and every later winner's
        # thread holds >= 8 channel-0 pixels that fired before it
        o = o_out[0]
        fired = o.any(axis=0)
        chan = np.where(fired.any(axis=0), fired.argmax(axis=0), 0).reshape(-1)
        t0 = np.where(fired.any(axis=0), o.any(axis=1).argmax(axis=0), o.shape[0]).reshape(-1)
        hw = chan.size
        assert o_cnt[0] >= 3 and o_win[0, 0]["z"] == 0, o_cnt
        for k in range(1, o_cnt[0]):
            p = int(o_win[0, k]["y"]) * o.shape[3] + int(o_win[0, k]["x"])
            mine = np.arange(p % 1024, hw, 1024)
            assert ((chan[mine] == 0) & (t0[mine] < t0[p])).sum() >= 8, (i, k)
        assert np.array_equal(H(plan.spikes(out, i)), o_out[0]), i
        assert np.array_equal(H(plan.pooled(pooled, i)), o_pool[0]), i
        check_winners(wn, o_win, o_cnt, i)
        o_k = orc.convstdp(x[None], o_k, o_out, [(0.004, -0.003)], o_win, o_cnt, 1, pad)
    assert np.array_equal(H(conv.kernel), o_k)


def test_region_precision_vs_fp64(dev, orc):
    n, t, cin, h, w, cout = 2, 12, 8, 48, 48, 64
    rng = np.random.default_rng(35)
    x = spikes(rng, (n, t, cin, h, w))
    kern = np.clip(rng.normal(0.8, 0.05, (cout, cin, 7, 7)), 0, 1).astype(np.float32)
    pad = (3, 3, 3, 3)
    pot64 = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad, f64=True).reshape(n, t, cout, h, w)
    th = float(np.quantile(pot64, 0.7))
    conv = make_conv(dev, kern, pad)
    pots = torch.empty((n, t, cout, h, w), dtype=torch.float32, device=dev)
    spk.layer_step(conv, T(x, dev), th, 5, 10, learn=False, potentials=pots)
    p16 = H(pots).astype(np.float64)
    assert np.abs(p16 - pot64).max() <= cin * 49 * 2.0 ** -11 * kern.max()
    flips = np.mean((p16 > th) != (pot64 > th))
    assert flips <= 1e-3, flips
    assert np.array_equal(H(conv.kernel), kern)


def test_region_graph_replay(dev, orc):
    """Eager runs and replays of a captured run give the same results, step
    after step (the weights change in place each step)."""
    rng = np.random.default_rng(36)
    shapes = [(6, 24, 24), (8, 32, 28), (5, 20, 36)]
    xs = [spikes(rng, (t, 4, h, w)) for t, h, w in shapes]
    kern = np.clip(rng.normal(0.8, 0.05, (32, 4, 5, 5)), 0, 1).astype(np.float32)
    conv = make_conv(dev, kern, 2)
    plan = spk.LayerPlan(conv, shapes, 6.0, 2, 4, 2, 2, tune=True)
    packed = plan.pack([T(x, dev) for x in xs])
    bufs = plan.buffers()
    eager = []
    for _ in range(3):
        out, pooled, wn = plan.run(packed, buffers=bufs)
        eager.append((out.clone(), pooled.clone(), wn.data.clone(), wn.counts.clone(), conv.kernel.clone()))
    conv.kernel.copy_(T(kern, dev))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(packed, buffers=bufs)  # warm on the capture stream (weights step once)
    torch.cuda.current_stream().wait_stream(side)
    conv.kernel.copy_(T(kern, dev))
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(packed, buffers=bufs)
    conv.kernel.copy_(T(kern, dev))
    torch.cuda.synchronize()
    for e in eager:
        g.replay()
        torch.cuda.synchronize()
        assert e[3].sum() > 0
        assert torch.equal(bufs[0], e[0]) and torch.equal(bufs[1], e[1])
        assert torch.equal(bufs[2], e[2]) and torch.equal(bufs[3], e[3])
        assert torch.equal(conv.kernel, e[4])
