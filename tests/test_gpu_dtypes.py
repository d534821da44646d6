"""F16 and F64 values on the plain path, as the reference runs them
(opers.cpp:408-421: the conv's input converted to the kernel's type, the
output in the kernel's type; inhibit, WTA, fire and STDP in the values' own
type, cpu/inhibit.cpp, cpu/wta.cpp, cpu/stdp.cpp:121-137).

Bar: bit-exact against the typed oracle (oracle/typed.py, itself pinned to
the reference's own CPU code) for every stage; the conv in the backend's
fixed order (FP32 fma chain rounded once to F16 for F16 kernels, FP64 fma
chain for F64 kernels)."""

import numpy as np
import pytest
import torch

import spyker_amd as spk
from oracle import typed

pytestmark = pytest.mark.gpu

DT = {np.float16: torch.float16, np.float64: torch.float64}


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


def winners_np(w):
    d, c = H(w.data), H(w.counts)
    out = np.zeros(d.shape[:2], dtype=typed.WINNER)
    for i, f in enumerate("ctzyx"):
        out[f] = d[:, :, i]
    return out, c


def same_winners(w, c, gw, gc):
    assert np.array_equal(c, gc)
    for i in range(len(c)):
        assert np.array_equal(w[i, :c[i]], gw[i, :gc[i]])


def potentials(rng, shape, dt):
    x = rng.choice([0.0, 0.5, 1.25, 2.0, 2.0009765625, 3.0, 4.5, 2.2, 0.7], size=shape)
    x = x + rng.normal(0, 0.3, size=shape) * (rng.random(shape) < 0.5)
    return np.cumsum(x, axis=1).astype(dt)


def spikes(rng, shape, p=0.3):
    return np.cumsum(rng.random(shape) < p, axis=1).clip(0, 1).astype(np.uint8)


@pytest.mark.parametrize("cout", [37, 6])
@pytest.mark.parametrize("dt", [np.float16, np.float64])
@pytest.mark.parametrize("src", [np.uint8, np.float32, "own"])
def test_conv_kernel_types(dev, orc, dt, src, cout):
    rng = np.random.default_rng(71)
    n, t, cin, h, w = 2, 3, 5, 11, 13
    if src == np.uint8:
        x = spikes(rng, (n, t, cin, h, w))
    else:
        x = rng.normal(0, 1, (n, t, cin, h, w)).astype(dt if src == "own" else src)
    k = rng.normal(0.5, 0.3, (cout, cin, 3, 3)).astype(dt)
    got = H(spk.conv(T(x, dev), T(k, dev), 1, (1, 2, 0, 1)))
    want = typed.conv(orc, x.reshape(n * t, cin, h, w), k, 1, (1, 2, 0, 1))
    assert got.dtype == dt
    assert np.array_equal(got.reshape(want.shape), want)


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_threshold_fire_inhibit(dev, dt):
    rng = np.random.default_rng(72)
    x = potentials(rng, (3, 7, 6, 9, 10), dt)
    assert np.array_equal(H(spk.threshold(T(x, dev), 2.0)), typed.threshold(x, 2.0))
    assert np.array_equal(H(spk.fire(T(x, dev), 2.0)), typed.fire(x, 2.0))
    got, want = H(spk.inhibit(T(x, dev), 2.0)), typed.inhibit(x, 2.0)
    assert np.array_equal(got, want) and np.array_equal(np.signbit(got), np.signbit(want))


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_convwta_fcwta(dev, dt):
    rng = np.random.default_rng(73)
    x = potentials(rng, (3, 7, 6, 9, 10), dt)
    for r, k, th in [(1, 4, 0.0), ((2, 1), 9, 3.0), (0, 50, 100.0)]:
        gw, gc = winners_np(spk.convwta(T(x, dev), r, k, th))
        w, c = typed.convwta(x, r, k, th)
        same_winners(w, c, gw, gc)
        assert th > 10 or c.sum() > 0
    f = x.reshape(3, 7, -1)
    gw, gc = winners_np(spk.fcwta(T(f, dev), 3, 6, 1.0))
    w, c = typed.fcwta(f, 3, 6, 1.0)
    same_winners(w, c, gw, gc)


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_convstdp_kernel_types(dev, dt):
    rng = np.random.default_rng(74)
    n, t, cin, h, w, cout = 3, 6, 4, 10, 11, 9
    si = spikes(rng, (n, t, cin, h, w))
    so = spikes(rng, (n, t, cout, h, w), 0.15)
    kern = rng.uniform(0, 1, (cout, cin, 3, 3)).astype(dt)
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev, dtype="f16" if dt == np.float16 else "f64")
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003), spk.STDPConfig(0.02, -0.01, False, 0.2, 0.8)]
    win = np.zeros((n, 4), typed.WINNER)
    rows = [[(0, 0, 1, 2, 3), (1, 0, 4, 0, 10), (0, 0, 8, 9, 0)], [(1, 0, 1, 5, 5)], [(0, 0, 0, 0, 0), (0, 0, 1, 2, 3)]]
    cnt = np.array([len(r) for r in rows])
    for i, r in enumerate(rows):
        for j, e in enumerate(r):
            win[i, j] = e
    data = torch.zeros((n, 4, 5), dtype=torch.int64, device=dev)
    for i, f in enumerate("ctzyx"):
        data[:, :, i] = T(win[f], dev)
    conv.stdp(T(si, dev), spk.Winners(data, T(cnt, dev)), T(so, dev))
    cfg = [(0.004, -0.003), (0.02, -0.01, False, 0.2, 0.8)]
    want = typed.convstdp(si, kern, so, cfg, win, cnt, 1, (1, 1, 1, 1))
    assert np.array_equal(H(conv.kernel), want) and not np.array_equal(want, kern)


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_layer_step_in_the_kernel_type(dev, orc, dt):
    """layer_step with F16 or F64 weights runs the plain path in that type:
    conv, threshold, inhibit, convwta, fire, STDP and pool against the typed
    oracle, stage by stage."""
    rng = np.random.default_rng(75)
    n, t, cin, h, w, cout = 3, 8, 4, 12, 12, 16
    x = spikes(rng, (n, t, cin, h, w))
    kern = rng.normal(0.8, 0.05, (cout, cin, 3, 3)).clip(0, 1).astype(dt)
    conv = spk.Conv(cin, cout, 3, pad=1, device=dev, dtype="f16" if dt == np.float16 else "f64")
    conv.kernel = T(kern, dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    pots = torch.empty((n, t, cout, h, w), dtype=DT[dt], device=dev)
    out, wn, pooled = spk.layer_step(conv, T(x, dev), 6.0, 1, 5, 2, 2, potentials=pots)
    pot = typed.conv(orc, x.reshape(n * t, cin, h, w), kern, 1, (1, 1, 1, 1)).reshape(n, t, cout, h, w)
    assert np.array_equal(H(pots), pot)
    pot = typed.inhibit(typed.threshold(pot, 6.0))
    w_, c_ = typed.convwta(pot, 1, 5)
    assert c_.sum() > 0
    same_winners(w_, c_, *winners_np(wn))
    o = typed.fire(pot)
    assert np.array_equal(H(out), o)
    assert np.array_equal(H(pooled), orc.pool(o, 2, 2))
    assert np.array_equal(H(conv.kernel), typed.convstdp(x, kern, o, [(0.004, -0.003)], w_, c_, 1, (1, 1, 1, 1)))


def test_f64_to_f16_rounds_once(dev):
    """A double just above a half-way point between two halves rounds up;
    through F32 first it would land on the half-way point and round to even."""
    base = np.float16(1.0)
    ulp = 2.0 ** -10
    vals = np.array([1.0 + ulp / 2 + 2.0 ** -40, 1.0 + ulp / 2 - 2.0 ** -40, 1.0 + 1.5 * ulp + 2.0 ** -40,
                     -(1.0 + ulp / 2 + 2.0 ** -40), 3.0 + 2 * ulp * 1.5 + 2.0 ** -38], np.float64)
    want = vals.astype(np.float16)
    assert want[0] != base  # the single rounding goes up
    out = torch.empty(vals.shape, dtype=torch.float16, device=dev)
    x = T(vals, dev)
    spk.call("spk_cast", spk._ptr(x), spk._lib.F64, spk._ptr(out), spk._lib.F16, vals.size, spk._stream(x))
    assert np.array_equal(H(out), want)


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_pool_pad(dev, dt):
    rng = np.random.default_rng(76)
    x = np.where(rng.random((2, 3, 4, 9, 11)) < 0.3, 0.0, rng.normal(0, 1, (2, 3, 4, 9, 11)))
    x = np.where(rng.random(x.shape) < 0.1, -0.0, x).astype(dt)
    for k, s, p in [(2, 2, (0, 0, 0, 0)), (3, 2, (1, 0, 2, 1))]:
        got, want = H(spk.pool(T(x, dev), k, s, p)), typed.pool(x, k, s, p)
        assert np.array_equal(got, want) and np.array_equal(np.signbit(got), np.signbit(want))
    assert np.array_equal(H(spk.pad(T(x, dev), (2, 0, 1, 3), -1.5)), typed.pad(x, (2, 0, 1, 3), -1.5))
