"""The workloads' own layer shapes, end to end, against the oracle.

Covers what the smaller parity tests leave out:
  * the precision variant's layer: 128x128 bar images through the Gabor bank
    (4 orientations x 2 scales), rank code T=30, conv 8->64 7x7 pad 3,
    threshold 20, inhibit, convwta r5 k10, pool 2x2, STDP -- bit-exact to
    the oracle on dyadic weights, through spk_stdp_layer_fused and through
    every candidate tiling of a layer plan (all of them cover the 30 steps
    in more than one pass of the FP16 windows);
  * the batching variant's extremes: T in {24, 32}, S in {96, 256}, with
    every candidate tiling of the plan, including a 12-pass time loop;
  * threshold-flip rate <= 0.1 % and the potential bound
    K * 2^-11 * max|w| against FP64 on weights the precision tests above do
    not use: uniform U(0, 1), bimodal Beta(0.2, 0.2) and weights driven
    apart by 30 STDP steps -- at the layer-2 and the 128x128 shapes.
"""

import numpy as np
import pytest
import torch

import spyker_amd as spk
from spyker_amd import synthetic

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


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


def oracle_step(orc, x, k, th, radius, count, pad, pool):
    """The reference's step: conv, threshold, inhibit, convwta, fire, STDP, pool."""
    n, t, cin, h, w = x.shape
    cout, _, kh, kw = k.shape
    pt, pl, pb, pr = pad
    ho, wo = h + pt + pb - kh + 1, w + pl + pr - kw + 1
    pot = orc.conv(x.reshape(n * t, cin, h, w), k, 1, pad).reshape(n, t, cout, ho, wo)
    pot = orc.inhibit(orc.threshold(pot, th))
    win, cnt = orc.convwta(pot, radius, count)
    out = orc.fire(pot)
    del pot
    k2 = orc.convstdp(x, k, out, [(0.004, -0.003)], win, cnt, 1, pad)
    pooled = orc.pool(out, pool[0], pool[1]) if pool else None
    return out, win, cnt, k2, pooled


# -- the precision variant's layer -------------------------------------------

C3_S, C3_T = 128, 30
GABOR_BANKS = ((3, 4.0), (5, 8.0))  # (half size, wavelength): 7x7 lambda 4, 11x11 lambda 8


def c3_spikes(dev, orc, n, seed=3):
    """Bar images -> Gabor 4 orientations x 2 scales -> threshold 0 -> rank code T=30,
    on the device, each stage checked against the oracle."""
    img = synthetic.bars(n, C3_S, seed=seed).reshape(n, 1, C3_S, C3_S)
    x = T(img, dev)
    banks, obanks = [], []
    for size, lam in GABOR_BANKS:
        filt = [spk.GaborFilter(0.56 * lam, th, 0.5, lam, 0.0) for th in (0.0, 45.0, 90.0, 135.0)]
        gab = spk.Gabor(size, filt, pad=size, device=dev)
        banks.append(gab(x))
        obanks.append(orc.gabor(img, H(gab.kernel), (size,) * 4))
    g = torch.cat(banks, dim=1).contiguous()
    og = np.concatenate(obanks, axis=1)
    assert np.array_equal(H(g), og)
    spk.threshold(g, 0.0)
    og = orc.threshold(og, 0.0)
    s = spk.code(g, C3_T)
    os_ = orc.rank_code(og.reshape(n, -1), C3_T).reshape(n, C3_T, 8, C3_S, C3_S)
    assert np.array_equal(H(s), os_)
    return s, os_


def test_precision_variant_layer_exact(dev, orc):
    """Two full 128x128 samples of the precision variant's layer, T=30, on dyadic weights."""
    n, cout, k, pad, th = 2, 64, 7, (3, 3, 3, 3), 20.0
    xs, x = c3_spikes(dev, orc, n)
    rng = np.random.default_rng(301)
    kern = dyadic(rng, (cout, 8, k, k))
    o_out, o_win, o_cnt, o_k, o_pool = oracle_step(orc, x, kern, th, 5, 10, pad, (2, 2))
    assert o_cnt.min() > 0

    # through the entry point (samples too large for the LDS take the region path)
    conv = spk.Conv(8, cout, k, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    out, wn, pooled = spk.layer_step(conv, xs, th, 5, 10, 2, 2)
    assert np.array_equal(H(out), o_out)
    assert np.array_equal(H(pooled), o_pool)
    check_winners(wn, o_win, o_cnt)
    assert np.array_equal(H(conv.kernel), o_k)
    del out, pooled

    # through a plan, every candidate tiling
    plan = spk.LayerPlan(conv, [(C3_T, C3_S, C3_S)] * n, th, 5, 10, 2, 2, tune=True)
    cands, _, _ = plan.candidates()
    assert len(cands) >= 2 and all(ch >= 2 for _, ch in cands), cands
    bufs = plan.buffers()
    for i in range(len(cands)):
        for writes in (1, 0):  # outputs written by the region kernel, then by the separate writers
            conv.kernel.copy_(T(kern, dev))
            plan.select(i, writes)
            out, pooled, wn = plan.run(xs.reshape(-1), buffers=bufs)
            assert np.array_equal(H(out).reshape(o_out.shape), o_out), (cands[i], writes)
            assert np.array_equal(H(pooled).reshape(o_pool.shape), o_pool), (cands[i], writes)
            check_winners(wn, o_win, o_cnt)
            assert np.array_equal(H(conv.kernel), o_k), (cands[i], writes)


# -- the batching variant's extremes -----------------------------------------

def c4_spikes(dev, orc, shapes, seed=4):
    """Stroke images -> DoG on/off pairs at 3x3 and 7x7 (4 channels) -> threshold 0.01 -> rank code,
    per instance, on the device and checked against the oracle."""
    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    dev_x, host_x = [], []
    for i, (t, h, w) in enumerate(shapes):
        img = (synthetic.strokes(1, h, seed=seed + i).astype(np.float32) / 255).reshape(1, 1, h, w)
        f = torch.cat([d3(T(img, dev)), d7(T(img, dev))], dim=1).contiguous()
        of = np.concatenate([orc.dog(img, H(d3.kernel), (1,) * 4), orc.dog(img, H(d7.kernel), (3,) * 4)], axis=1)
        assert np.array_equal(H(f), of)
        spk.threshold(f, 0.01)
        s = spk.code(f, t)
        os_ = orc.rank_code(orc.threshold(of, 0.01).reshape(1, -1), t).reshape(1, t, 4, h, w)
        assert np.array_equal(H(s), os_)
        dev_x.append(s[0])
        host_x.append(os_)
    return dev_x, host_x


@pytest.mark.parametrize("shapes", [
    [(32, 256, 256), (24, 96, 96), (32, 96, 96)],  # tmax 32: the tallest region covers it in 2 passes
    [(24, 256, 256), (24, 96, 96), (8, 32, 32)],   # tmax 24: the tallest region in 12 passes
])
def test_batching_variant_long_steps(dev, orc, shapes):
    cin, cout, k, pad, th, radius, count, pool = 4, 32, 5, (2, 2, 2, 2), 12.0, 2, 4, (2, 2)
    xd, xh = c4_spikes(dev, orc, shapes)
    rng = np.random.default_rng(401)
    kern = dyadic(rng, (cout, cin, k, k))
    conv = spk.Conv(cin, cout, k, pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    plan = spk.LayerPlan(conv, shapes, th, radius, count, *pool, tune=True)
    cands, _, _ = plan.candidates()
    assert max(ch for _, ch in cands) >= 2, cands
    packed = plan.pack(xd)
    # the oracle: every instance's forward pass on the original weights, then
    # STDP instance by instance in order
    ref = []
    o_k = kern
    for x in xh:
        o_out, o_win, o_cnt, _, o_pool = oracle_step(orc, x, kern, th, radius, count, pad, pool)
        o_k = orc.convstdp(x, o_k, o_out, [(0.004, -0.003)], o_win, o_cnt, 1, pad)
        ref.append((o_out, o_win, o_cnt, o_pool))
    assert sum(r[2].sum() for r in ref) > 0
    bufs = plan.buffers()
    for c in range(len(cands)):
        for writes in (1, 0):
            conv.kernel.copy_(T(kern, dev))
            plan.select(c, writes)
            out, pooled, wn = plan.run(packed, buffers=bufs)
            for i, (o_out, o_win, o_cnt, o_pool) in enumerate(ref):
                assert np.array_equal(H(plan.spikes(out, i)), o_out[0]), (cands[c], writes, i)
                assert np.array_equal(H(plan.pooled(pooled, i)), o_pool[0]), (cands[c], writes, i)
                check_winners(wn, o_win, o_cnt, i)
            assert np.array_equal(H(conv.kernel), o_k), (cands[c], writes)


# -- precision on weights away from N(0.8, 0.05) -------------------------------

def flip_check(dev, orc, x, kern, pad, th_q):
    """FP16-path potentials against FP64: the error bound and the flip rate."""
    n, t, cin, h, w = x.shape
    cout, _, kh, kw = kern.shape
    ho, wo = h + pad[0] + pad[2] - kh + 1, w + pad[1] + pad[3] - kw + 1
    p64 = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, pad, f64=True).reshape(n, t, cout, ho, wo)
    th = float(np.quantile(p64[p64 > 0], th_q))
    conv = spk.Conv(cin, cout, (kh, kw), pad=pad, device=dev)
    conv.kernel = T(kern, dev)
    pots = torch.empty((n, t, cout, ho, wo), dtype=torch.float32, device=dev)
    spk.layer_step(conv, T(x, dev), th, 1, 4, learn=False, potentials=pots)
    p16 = H(pots).astype(np.float64)
    err = np.abs(p16 - p64).max()
    assert err <= cin * kh * kw * 2.0 ** -11 * np.abs(kern).max(), err
    flips = np.mean((p16 > th) != (p64 > th))
    assert flips <= 1e-3, flips
    return flips, err


def weights(kind, rng, shape):
    if kind == "uniform":
        return rng.random(shape).astype(np.float32)
    return rng.beta(0.2, 0.2, shape).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "bimodal"])
def test_flip_rate_layer2_shape(dev, orc, kind):
    rng = np.random.default_rng(501)
    n, t, cin, h, w, cout = 8, 15, 30, 14, 14, 250
    x = np.maximum.accumulate((rng.random((n, t, cin, h, w)) < 0.1).astype(np.uint8), axis=1)
    flip_check(dev, orc, x, weights(kind, rng, (cout, cin, 3, 3)), (1, 1, 1, 1), 0.7)


@pytest.mark.parametrize("kind", ["uniform", "bimodal"])
def test_flip_rate_precision_variant_shape(dev, orc, kind):
    xs, x = c3_spikes(dev, orc, 1, seed=13)
    rng = np.random.default_rng(502)
    flip_check(dev, orc, x, weights(kind, rng, (64, 8, 7, 7)), (3, 3, 3, 3), 0.7)


def test_flip_rate_after_stdp_training(dev, orc):
    """Weights trained apart by 30 fused STDP steps (fast rates, so they reach
    the bounds): the FP16 path still flips <= 0.1 % of thresholds."""
    rng = np.random.default_rng(503)
    n, t, cin, h, w, cout = 64, 15, 30, 14, 14, 250
    x = np.maximum.accumulate((rng.random((n, t, cin, h, w)) < 0.1).astype(np.uint8), axis=1)
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev)
    conv.kernel = T(np.clip(rng.normal(0.8, 0.05, (cout, cin, 3, 3)), 0, 1).astype(np.float32), dev)
    conv.stdpconfig = [spk.STDPConfig(0.2, -0.15)]
    xd = T(x, dev)
    for _ in range(30):
        spk.layer_step(conv, xd, 10.0, 1, 8)
    kern = H(conv.kernel)
    # trained apart: a good share of the weights sit near the bounds
    assert np.mean((kern < 0.1) | (kern > 0.9)) > 0.2
    flip_check(dev, orc, x[:8], kern, (1, 1, 1, 1), 0.7)
