"""Parity of every device op with the oracle and the reference vectors.

Bar: bit-exact (integer, byte and index outputs, and the FP32 plain path's
floating outputs, whose order of operations the oracle fixes).  All calls go
through the C ABI via the spyker_amd front end.
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


def winners_np(w):
    d, c = H(w.data), H(w.counts)
    out = np.zeros(d.shape[:2], dtype=[("c", "<i8"), ("t", "<i8"), ("z", "<i8"), ("y", "<i8"), ("x", "<i8")])
    for i, f in enumerate("ctzyx"):
        out[f] = d[:, :, i]
    return out, c


def same_winners(w, c, gw, gc):
    assert np.array_equal(c, gc)
    for i in range(len(c)):
        assert np.array_equal(w[i, :c[i]], gw[i, :gc[i]])


def test_threshold_fire(dev, golden):
    p = golden["pot"]
    assert np.array_equal(H(spk.threshold(T(p, dev), 15.0)), golden["threshold_15"])
    assert np.array_equal(H(spk.threshold(T(p, dev), 15.0, 3.0)), golden["threshold_15_v3"])
    assert np.array_equal(H(spk.fire(T(p, dev), 15.0)), golden["fire_15"])
    assert np.array_equal(H(spk.fire(T(p, dev), 0.0)), golden["fire_0"])
    rng = np.random.default_rng(0)
    for n in (1, 15, 16, 17, 4099, 1 << 20):
        x = (rng.random(n) * 30).astype(np.float32)
        assert np.array_equal(H(spk.threshold(T(x, dev), 15.0)), np.where(x <= 15, 0, x).astype(np.float32))
        assert np.array_equal(H(spk.fire(T(x, dev), 15.0)), (x > 15).astype(np.uint8))
    x = (rng.random(1001) * 30).astype(np.float16)
    assert np.array_equal(H(spk.threshold(T(x, dev), 15.0)), np.where(x <= np.float16(15), 0, x).astype(np.float16))


def test_inhibit(dev, golden, orc):
    assert np.array_equal(H(spk.inhibit(T(golden["threshold_15"], dev))), golden["inhibit"])
    assert np.array_equal(H(spk.inhibit(T(golden["threshold_15"], dev), 5.0)), golden["inhibit_th5"])
    rng = np.random.default_rng(1)
    p = orc.threshold(np.floor(rng.random((5, 15, 30, 14, 14)) * 60).astype(np.float32) / 2, 20.0)
    assert np.array_equal(H(spk.inhibit(T(p, dev))), orc.inhibit(p))


def test_convwta(dev, golden, orc):
    inh = golden["inhibit"]
    for name, (r, k) in {"r1k4": (1, 4), "r3k5": (3, 5), "r12k3": ((1, 2), 3), "r0k20": (0, 20)}.items():
        w, c = winners_np(spk.convwta(T(inh, dev), r, k))
        same_winners(w, c, golden["convwta_%s_w" % name], golden["convwta_%s_c" % name])
    w, c = winners_np(spk.convwta(T(golden["threshold_15"], dev), 1, 6, 16.0))
    same_winners(w, c, golden["convwta_th16_w"], golden["convwta_th16_c"])
    w, c = winners_np(spk.convwta(T(golden["spikes_u8"], dev), 2, 5))
    same_winners(w, c, golden["convwta_u8_w"], golden["convwta_u8_c"])
    rng = np.random.default_rng(2)
    p = orc.inhibit(orc.threshold(rng.random((7, 15, 40, 14, 14)).astype(np.float32) * 30, 24.0))
    w, c = winners_np(spk.convwta(T(p, dev), 1, 8))
    ow, oc = orc.convwta(p, 1, 8)
    same_winners(w, c, ow, oc)


def test_fcwta(dev, golden):
    w, c = winners_np(spk.fcwta(T(golden["fc_pot"], dev), 2, 6))
    same_winners(w, c, golden["fcwta_w"], golden["fcwta_c"])


@pytest.mark.parametrize("pad", [(0, 0, 0, 0), (1, 1, 1, 1), (1, 0, 2, 1)])
def test_pool(dev, golden, pad):
    key = "".join(str(q) for q in pad)
    s8, p = golden["spikes_u8"], golden["pot"] - 5.0
    assert np.array_equal(H(spk.pool(T(s8, dev), 2, 2, pad)), golden["pool_u8_k2s2_p" + key])
    assert np.array_equal(H(spk.pool(T(s8, dev), 3, 3, pad)), golden["pool_u8_k3s3_p" + key])
    assert np.array_equal(H(spk.pool(T(p, dev), 3, 2, pad)), golden["pool_f32_k3s2_p" + key])


@pytest.mark.parametrize("kind,pad,value,key", [
    ("u8", (1, 1, 1, 1), 0, "pad_u8_p1111_v0"), ("u8", (2, 0, 3, 1), 7, "pad_u8_p2031_v7"),
    ("f32", (0, 1, 0, 2), 0.0, "pad_f32_p0102_v0"), ("f32", (3, 2, 1, 0), -1.5, "pad_f32_p3210_vm1p5")])
def test_pad(dev, golden, orc, kind, pad, value, key):
    """Dense pad (cuda_pad, impl.h:236) against the reference's cpu_pad vectors
    (cpu/pool.cpp:41-51), and a larger plane set against the oracle."""
    x = golden["pad_in_" + kind]
    got = H(spk.pad(T(x, dev), pad, value))
    assert np.array_equal(got.reshape(golden[key].shape), golden[key])
    rng = np.random.default_rng(9)
    big = (rng.integers(0, 256, (3, 5, 7, 33, 41)).astype(np.uint8) if kind == "u8"
           else rng.normal(0, 1, (3, 5, 7, 33, 41)).astype(np.float32))
    got = H(spk.pad(T(big, dev), pad, value))
    assert np.array_equal(got, orc.pad(big, pad, value))


@pytest.mark.parametrize("case", [
    # n, t, cin, h, w, cout, kh, kw, stride, pad
    (4, 3, 4, 28, 28, 30, 5, 5, 1, (2, 2, 2, 2)),
    (3, 5, 30, 14, 14, 250, 3, 3, 1, (1, 1, 1, 1)),
    (2, 2, 8, 40, 37, 20, 7, 7, 1, (3, 3, 3, 3)),
    (2, 2, 3, 17, 300, 5, 3, 5, 2, (0, 2, 1, 1)),
])
def test_conv_bit_exact(dev, orc, case):
    n, t, cin, h, w, cout, kh, kw, stride, pad = case
    rng = np.random.default_rng(3)
    x = np.maximum.accumulate((rng.random((n, t, cin, h, w)) < 0.2).astype(np.uint8), axis=1)
    k = np.clip(rng.normal(0.8, 0.05, (cout, cin, kh, kw)), 0, 1).astype(np.float32)
    got = H(spk.conv(T(x, dev), T(k, dev), stride, pad))
    want = orc.conv(x.reshape(n * t, cin, h, w), k, stride, pad)
    assert np.array_equal(got.reshape(want.shape), want)
    xf = rng.random((n * t, cin, h, w)).astype(np.float32)
    got = H(spk.conv(T(xf[:, None].reshape(n * t, 1, cin, h, w), dev), T(k, dev), stride, pad))
    assert np.array_equal(got.reshape(want.shape), orc.conv(xf, k, stride, pad))


def test_conv_f64_check_path(dev, orc):
    rng = np.random.default_rng(4)
    x = (rng.random((6, 30, 14, 14)) < 0.3).astype(np.uint8)
    k = rng.random((50, 30, 3, 3)).astype(np.float32)
    out = torch.empty((6, 50, 14, 14), dtype=torch.float64, device=dev)
    from spyker_amd import _lib
    xd, kd = T(x, dev), T(k, dev)
    spk.call("spk_conv", spk._ptr(xd), _lib.U8, spk._ptr(kd), _lib.F32, spk._ptr(out), _lib.F64, 6, 30, 14, 14, 50,
             3, 3, 1, 1, 1, 1, 1, 1, spk._stream(xd))
    assert np.array_equal(H(out), orc.conv(x, k, 1, (1, 1, 1, 1), f64=True))


def test_fc(dev, orc):
    rng = np.random.default_rng(5)
    x = (rng.random((3, 7, 300)) < 0.3).astype(np.uint8)
    k = rng.random((33, 300)).astype(np.float32)
    assert np.array_equal(H(spk.fc(T(x, dev), T(k, dev))), orc.fc(x, k))


def test_dog_gabor(dev, orc):
    rng = np.random.default_rng(6)
    img = rng.random((5, 1, 28, 28)).astype(np.float32)
    d = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    assert np.array_equal(H(d(T(img, dev))), orc.dog(img, H(d.kernel), (3, 3, 3, 3)))
    g = spk.Gabor(3, [spk.GaborFilter(2.24, th, 0.5, 4.0, 0.0) for th in (0, 45, 90, 135)], pad=3, device=dev)
    assert np.array_equal(H(g(T(img, dev))), orc.gabor(img, H(g.kernel), (3, 3, 3, 3)))
    u8 = (rng.random((2, 1, 28, 28)) * 255).astype(np.uint8)
    assert np.array_equal(H(d(T(u8, dev))), orc.dog(u8.astype(np.float32), H(d.kernel), (3, 3, 3, 3)))
    # up to 4 filters a block combines them in its epilogue; 5 take the two-pass form
    for nf in (1, 4, 5):
        dn = spk.DoG(2, [spk.DoGFilter(1 + 0.25 * i, 2 + 0.5 * i) for i in range(nf)], pad=(2, 1, 0, 2), device=dev)
        assert np.array_equal(H(dn(T(img, dev))), orc.dog(img, H(dn.kernel), (2, 1, 0, 2)))


def test_log(dev, orc):
    rng = np.random.default_rng(9)
    img = rng.random((4, 2, 28, 28)).astype(np.float32)
    lg = spk.LoG(3, [1.0, 2.0, 3.0], pad=(3, 2, 1, 3), device=dev)
    got = H(lg(T(img, dev)))
    assert got.shape == (4, 2 * 2 * 3, 28 + 4 - 6, 28 + 5 - 6)
    assert np.array_equal(got, orc.log(img, H(lg.kernel), (3, 2, 1, 3)))
    u8 = (rng.random((1, 28, 28)) * 255).astype(np.uint8)  # 3-D input is one sample (to4)
    assert np.array_equal(H(lg(T(u8, dev))), orc.log(u8.astype(np.float32)[None], H(lg.kernel), (3, 2, 1, 3)))
    for stds in ([1.5], [1.0, 1.5, 2.0, 2.5], [1.0, 1.25, 1.5, 1.75, 2.0]):
        ln = spk.LoG(2, stds, pad=2, device=dev)
        assert np.array_equal(H(ln(T(img, dev))), orc.log(img, H(ln.kernel), (2, 2, 2, 2)))


def test_rank_code(dev, golden, orc):
    x = golden["code_in"]
    for t in (1, 15, 30):
        got = H(spk.code(T(x, dev), t))
        assert np.array_equal(got, orc.rank_code(x, t))
        assert np.array_equal(H(spk.code(T(x[:4], dev), t, sort=False)), golden["code_latency_T%d" % t])
    rng = np.random.default_rng(7)
    for shape in [(3, 3136), (2, 2049), (2, 8 * 128 * 128), (1, 5000)]:
        y = rng.random(shape).astype(np.float32)
        y[y < 0.4] = 0
        y = np.round(y * 64) / 64  # many ties
        assert np.array_equal(H(spk.code(T(y, dev), 15)), orc.rank_code(y, 15))


def test_convstdp(dev, golden, orc):
    g = golden
    conv = spk.Conv(3, 4, 3, pad=1, device=dev)
    conv.kernel = T(g["stdp_kernel"], dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    wn = spk.Winners(T(np.stack([g["stdp_win"][f] for f in "ctzyx"], -1), dev), T(g["stdp_cnt"], dev))
    conv.stdp(T(g["stdp_in"], dev), wn, T(g["stdp_out"], dev))
    assert np.array_equal(H(conv.kernel), g["stdp_kernel_after"])
    conv.kernel = T(g["stdp_kernel"], dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003), spk.STDPConfig(-0.002, 0.005, False, 0.1, 0.9)]
    wr = spk.Winners(T(np.stack([g["stdp_win_rstdp"][f] for f in "ctzyx"], -1), dev), T(g["stdp_cnt"], dev))
    conv.stdp(T(g["stdp_in"], dev), wr, T(g["stdp_out"], dev))
    assert np.array_equal(H(conv.kernel), g["stdp_kernel_after_rstdp"])
    # a configuration index out of range fails the call and leaves the weights alone
    conv.kernel = T(g["stdp_kernel"], dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    with pytest.raises(RuntimeError, match="configuration index"):
        conv.stdp(T(g["stdp_in"], dev), wr, T(g["stdp_out"], dev))
    assert np.array_equal(H(conv.kernel), g["stdp_kernel"])


def test_fcstdp(dev, golden, orc):
    g = golden
    fc = spk.FC(40, 10, device=dev)
    fc.kernel = T(g["fcstdp_kernel"], dev)
    fc.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    wn = spk.Winners(T(np.stack([g["fcstdp_win"][f] for f in "ctzyx"], -1), dev), T(g["fcstdp_cnt"], dev))
    fc.stdp(T(g["fcstdp_in"], dev), wn, T(g["fcstdp_out"], dev))
    want = orc.fcstdp(g["fcstdp_in"], g["fcstdp_kernel"], g["fcstdp_out"], [(0.004, -0.003)], g["fcstdp_win"],
                      g["fcstdp_cnt"])
    assert np.array_equal(H(fc.kernel), want)


def test_layer_reference_vector(dev, golden):
    """The reference's own pipeline vector, stage by stage on the device."""
    g = golden
    conv = spk.Conv(1, 8, 5, pad=2, device=dev)
    conv.kernel = T(g["layer_kernel"], dev)
    conv.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    x = T(g["layer_in"], dev)
    pot = spk.inhibit(spk.threshold(conv(x), 8.0))
    assert np.array_equal(H(pot), g["layer_pot"])
    win = spk.convwta(pot, 2, 5)
    w, c = winners_np(win)
    same_winners(w, c, g["layer_win"], g["layer_cnt"])
    out = spk.fire(pot)
    assert np.array_equal(H(out), g["layer_out"])
    conv.stdp(x, win, out)
    assert np.array_equal(H(conv.kernel), g["layer_kernel_after"])
    assert np.array_equal(H(spk.pool(out, 2)), g["layer_pool"])


def test_readout_gather_scatter(dev, golden, orc):
    """rank/rate gather and rank scatter (cpu/time.cpp) against the reference vectors and the
    oracle, over input and output types."""
    rd, rf = golden["readout_in_u8"], golden["readout_in_f32"]
    for th in (0, 2):
        for dt in ("u8", "u16", "i32", "f32"):
            got = H(spk.gather(T(rd, dev), th, dt)).astype(np.int64)
            assert np.array_equal(got, golden["rank_gather_u8_th%d" % th]), (th, dt)
    assert np.array_equal(H(spk.gather(T(rf, dev), 0.5, "i32")), golden["rank_gather_f32_th05"])
    assert np.array_equal(H(spk.gather(T(rf, dev), code="rate", dtype="f32")), golden["rate_gather_f32"])
    tu, tf = golden["scatter_in_u8"], golden["scatter_in_f32"]
    assert np.array_equal(H(spk.scatter(T(tu, dev), 10)), golden["rank_scatter_u8_T10"])
    assert np.array_equal(H(spk.scatter(T(tf, dev), 10)), golden["rank_scatter_f32_T10"])
    assert np.array_equal(H(spk.scatter(T(tf, dev), 10, "f32")), golden["rank_scatter_f32_T10"].astype(np.float32))
    # larger, and a round trip through the layer's own spikes
    rng = np.random.default_rng(41)
    s = np.maximum.accumulate((rng.random((16, 15, 30, 14, 14)) < 0.05).astype(np.uint8), axis=1)
    first = H(spk.gather(T(s, dev)))
    assert np.array_equal(first.astype(np.int32), orc.rank_gather(s))
    assert np.array_equal(H(spk.scatter(T(first, dev), 15)), s)
    neg = np.array([[-3, 0, 2, 20]], np.int32)  # negative times compare as large unsigned values
    assert np.array_equal(H(spk.scatter(T(neg, dev), 4))[0, :, 0], np.zeros(4, np.uint8))


def test_two_layer_forward_features(dev, orc):
    """The inference forward the examples feed to a classifier
    (examples/rezghi_mnist_dense.py:82-85): pool(fire(conv1(x), th1), 2),
    pool(fire(conv2(.), th2), 3), then T - gather(.), flattened; equal to the
    oracle's composition of the same reference functions."""
    from spyker_amd import synthetic
    n, t = 6, 15
    img = synthetic.strokes(n, 28, seed=43).astype(np.float32) / 255
    d3 = spk.DoG(1, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=1, device=dev)
    d7 = spk.DoG(3, [spk.DoGFilter(1, 2), spk.DoGFilter(2, 1)], pad=3, device=dev)
    xi = T(img, dev).reshape(n, 1, 28, 28)
    f = torch.cat([d3(xi), d7(xi)], dim=1).contiguous()
    spk.threshold(f, 0.01)
    x = spk.code(f, t)
    rng = np.random.default_rng(44)
    k1 = np.clip(rng.normal(0.5, 0.02, (20, 4, 5, 5)), 0, 1).astype(np.float32)
    k2 = np.clip(rng.normal(0.5, 0.02, (40, 20, 3, 3)), 0, 1).astype(np.float32)
    c1, c2 = spk.Conv(4, 20, 5, pad=2, device=dev), spk.Conv(20, 40, 3, pad=1, device=dev)
    c1.kernel, c2.kernel = T(k1, dev), T(k2, dev)
    a = spk.pool(spk.fire(c1(x), 16.0), 2)
    a = spk.pool(spk.fire(c2(a), 5.0), 3)
    feats = (t - spk.gather(a).to(torch.int32)).flatten(1)
    xo = H(x)
    o = orc.conv(xo.reshape(n * t, 4, 28, 28), k1, 1, (2, 2, 2, 2)).reshape(n, t, 20, 28, 28)
    o = orc.pool(orc.fire(o, 16.0), 2, 2)
    o = orc.conv(o.reshape(n * t, 20, 14, 14), k2, 1, (1, 1, 1, 1)).reshape(n, t, 40, 14, 14)
    o = orc.pool(orc.fire(o, 5.0), 3, 3)
    want = (t - orc.rank_gather(o)).reshape(n, -1)
    assert 0 < (want > 0).mean() < 1
    assert np.array_equal(H(feats), want)


# spk_type ordinal -> numpy type (the reference's Type order, src/spyker/utils.h:96-117)
ALL_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64, np.float16,
             np.float32, np.float64]


def _raw(a, dev):
    """A device byte buffer holding numpy array a."""
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(dev)


def test_cast_and_fill_every_type(dev):
    """cuda_cast / cuda_fill for every element type (cpu/memory.cpp:54-70: std::copy's
    conversion and std::fill), on values every type holds exactly."""
    vals = np.array([0, 1, 2, 3, 7, 100, 120, 5, 64, 127, 9, 33, 17], np.float64)
    s = torch.cuda.current_stream().cuda_stream
    for i, ti in enumerate(ALL_TYPES):
        src = _raw(vals.astype(ti), dev)
        for o, to in enumerate(ALL_TYPES):
            dst = torch.zeros(len(vals) * np.dtype(to).itemsize, dtype=torch.uint8, device=dev)
            spk.call("spk_cast", spk._ptr(src), i, spk._ptr(dst), o, len(vals), s)
            got = H(dst).view(to)
            assert np.array_equal(got, vals.astype(ti).astype(to)), (ti, to)
        buf = torch.zeros(1000 * np.dtype(ti).itemsize, dtype=torch.uint8, device=dev)
        spk.call("spk_fill", spk._ptr(buf), i, 999, 42.0, s)
        got = H(buf).view(ti)
        assert (got[:999] == ti(42)).all() and got[999] == 0, ti
