"""Pins the CPU oracle (oracle/spyker_oracle.c) before anything is checked against it.

What pins each oracle function:

  rank_code (sort, latency)   reference vectors (the reference's own cpu/code.cpp,
                              compiled in oracle/_ref) incl. ties and a constant
                              sample; cumulative-spike invariant
  threshold, fire             reference vectors; numpy closed form
  inhibit                     reference vectors; one channel per pixel invariant
  convwta (f32, u8), fcwta    reference vectors; winner invariants
  pool (u8, f32, pads)        reference vectors; torch max_pool2d closed form
  convstdp, fcstdp            reference vectors (incl. R-STDP configs, stride 2,
                              asymmetric pad); only winner rows change, bounds
  stdp_update                 reference vectors through convstdp; exact rational
                              evaluation of the F64 formula on sampled weights
  conv f32/f64, fc            torch float64 conv2d/matmul; exact on dyadic
                              weights (orientation, pad order, stride)
  dog, gabor                  torch float64 conv2d of the same kernels
  gaussian/gabor kernels      reference vectors; unit sum

The reference's conv/FC arithmetic is oneDNN/cuDNN, absent here: the FP32
summation order of orc_conv_f32/orc_fc_f32 is this project's canonical order
and its parity with the reference is unpinned (only pinned to FP64 within
rounding).
"""

from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle


def tie_blocks_agree(x, a, b):
    """Spike times agree element by element except inside a block of equal
    values, where the reference's std::sort leaves the order unspecified:
    there the multiset of times must agree."""
    ta, tb = a.sum(axis=0), b.sum(axis=0)  # cumulative spikes: count = T - first spike time
    if np.array_equal(ta, tb):
        return True
    for v in np.unique(x[ta != tb]):
        m = x == v
        if m.sum() < 2 or not np.array_equal(np.sort(ta[m]), np.sort(tb[m])):
            return False
    return np.array_equal(ta[~np.isin(x, x[ta != tb])], tb[~np.isin(x, x[ta != tb])])


def test_rank_code_matches_reference_vectors(orc, golden):
    x = golden["code_in"]
    for t in (1, 15, 30):
        got, want = orc.rank_code(x, t, True), golden["code_sort_T%d" % t]
        for i in range(x.shape[0]):
            if i == 1:  # sample 1 holds a 40-way tie that straddles time boundaries
                assert tie_blocks_agree(x[i], got[i], want[i])
            else:
                assert np.array_equal(got[i], want[i])
        assert np.array_equal(orc.rank_code(x[:4], t, False), golden["code_latency_T%d" % t])


def test_rank_code_is_cumulative(orc):
    rng = np.random.default_rng(1)
    x = rng.random((3, 500)).astype(np.float32)
    s = orc.rank_code(x, 20)
    assert np.all(np.diff(s.astype(np.int16), axis=1) >= 0)
    # the largest value spikes at t=0, the minimum never spikes
    for i in range(3):
        assert s[i, 0, np.argmax(x[i])] == 1
        assert s[i, -1, np.argmin(x[i])] == 0


def test_threshold_fire(orc, golden):
    p = golden["pot"]
    assert np.array_equal(orc.threshold(p, 15.0), golden["threshold_15"])
    assert np.array_equal(orc.threshold(p, 15.0, 3.0), golden["threshold_15_v3"])
    assert np.array_equal(orc.fire(p, 15.0), golden["fire_15"])
    assert np.array_equal(orc.fire(p, 0.0), golden["fire_0"])
    assert np.array_equal(orc.threshold(p, 15.0), np.where(p <= 15.0, 0.0, p).astype(np.float32))
    assert np.array_equal(orc.fire(p, 15.0), (p > 15.0).astype(np.uint8))


def test_inhibit(orc, golden):
    th = golden["threshold_15"]
    out = orc.inhibit(th)
    assert np.array_equal(out, golden["inhibit"])
    assert np.array_equal(orc.inhibit(th, 5.0), golden["inhibit_th5"])
    nz = (out != 0).any(axis=1)  # [n, c, h, w]
    assert np.all(nz.sum(axis=1) <= 1)


def _check_winners(w, c, count, ry, rx):
    for i in range(len(c)):
        ws = w[i, :c[i]]
        assert len(set(ws["z"].tolist())) == len(ws)
        for a in range(len(ws)):
            for b in range(a):
                assert not (abs(ws[a]["y"] - ws[b]["y"]) <= ry and abs(ws[a]["x"] - ws[b]["x"]) <= rx)
        assert c[i] <= count


def test_convwta(orc, golden):
    inh = golden["inhibit"]
    for name, (r, k) in {"r1k4": (1, 4), "r3k5": (3, 5), "r12k3": ((1, 2), 3), "r0k20": (0, 20)}.items():
        w, c = orc.convwta(inh, r, k)
        assert np.array_equal(c, golden["convwta_%s_c" % name])
        assert np.array_equal(w, golden["convwta_%s_w" % name])
        ry, rx = (r, r) if np.isscalar(r) else r
        _check_winners(w, c, k, ry, rx)
    w, c = orc.convwta(golden["threshold_15"], 1, 6, 16.0)
    assert np.array_equal(c, golden["convwta_th16_c"]) and np.array_equal(w, golden["convwta_th16_w"])
    w, c = orc.convwta(golden["spikes_u8"], 2, 5)
    assert np.array_equal(c, golden["convwta_u8_c"]) and np.array_equal(w, golden["convwta_u8_w"])


def test_fcwta(orc, golden):
    w, c = orc.fcwta(golden["fc_pot"], 2, 6)
    assert np.array_equal(c, golden["fcwta_c"]) and np.array_equal(w, golden["fcwta_w"])


@pytest.mark.parametrize("pad", [(0, 0, 0, 0), (1, 1, 1, 1), (1, 0, 2, 1)])
def test_pool(orc, golden, pad):
    key = "".join(str(q) for q in pad)
    s8, p = golden["spikes_u8"], golden["pot"] - 5.0
    assert np.array_equal(orc.pool(s8, 2, 2, pad), golden["pool_u8_k2s2_p" + key])
    assert np.array_equal(orc.pool(s8, 3, 3, pad), golden["pool_u8_k3s3_p" + key])
    got = orc.pool(p, 3, 2, pad)
    assert np.array_equal(got, golden["pool_f32_k3s2_p" + key])
    # closed form: zero pad then max_pool2d
    pt, pl, pb, pr = pad
    tp = torch.nn.functional.pad(torch.from_numpy(p.reshape(-1, 9, 9)), (pl, pr, pt, pb))
    ref = torch.nn.functional.max_pool2d(tp[:, None], 3, 2)[:, 0].numpy()
    assert np.array_equal(got.reshape(ref.shape), ref)


PAD_CASES = [("u8", (1, 1, 1, 1), 0, "pad_u8_p1111_v0"), ("u8", (2, 0, 3, 1), 7, "pad_u8_p2031_v7"),
             ("f32", (0, 1, 0, 2), 0.0, "pad_f32_p0102_v0"), ("f32", (3, 2, 1, 0), -1.5, "pad_f32_p3210_vm1p5")]


@pytest.mark.parametrize("kind,pad,value,key", PAD_CASES)
def test_pad(orc, golden, kind, pad, value, key):
    x = golden["pad_in_" + kind]
    got = orc.pad(x, pad, value)
    assert np.array_equal(got, golden[key])
    # closed form: torch constant pad of the last two dimensions
    pt, pl, pb, pr = pad
    ref = torch.nn.functional.pad(torch.from_numpy(x.astype(np.float32)), (pl, pr, pt, pb), value=float(value))
    assert np.array_equal(got.astype(np.float32), ref.numpy())


def test_convstdp(orc, golden):
    g = golden
    k = orc.convstdp(g["stdp_in"], g["stdp_kernel"], g["stdp_out"], [(0.004, -0.003)], g["stdp_win"], g["stdp_cnt"],
                     1, (1, 1, 1, 1))
    assert np.array_equal(k, g["stdp_kernel_after"])
    changed = np.where((k != g["stdp_kernel"]).any(axis=(1, 2, 3)))[0]
    won = set(g["stdp_win"]["z"][np.arange(g["stdp_win"].shape[1])[None, :] < g["stdp_cnt"][:, None]].tolist())
    assert set(changed.tolist()) <= won
    assert k.min() >= 0 and k.max() <= 1
    k = orc.convstdp(g["stdp_in"], g["stdp_kernel"], g["stdp_out"], [(0.004, -0.003), (-0.002, 0.005, False, 0.1, 0.9)],
                     g["stdp_win_rstdp"], g["stdp_cnt"], 1, (1, 1, 1, 1))
    assert np.array_equal(k, g["stdp_kernel_after_rstdp"])
    w = g["stdp_win"].copy()
    w["y"], w["x"] = np.minimum(w["y"], 3), np.minimum(w["x"], 3)
    k = orc.convstdp(g["stdp_in"], g["stdp_kernel"], np.ascontiguousarray(g["stdp_out"][:, :, :, :4, :4]),
                     [(0.01, -0.02)], w, g["stdp_cnt"], 2, (1, 1, 0, 0))
    assert np.array_equal(k, g["stdp_kernel_after_s2"])


def test_convstdp_rejects_bad_config(orc, golden):
    g = golden
    w = g["stdp_win"].copy()
    w["c"] = 3
    with pytest.raises(RuntimeError):
        orc.convstdp(g["stdp_in"], g["stdp_kernel"], g["stdp_out"], [(0.004, -0.003)], w, g["stdp_cnt"], 1,
                     (1, 1, 1, 1))


def test_fcstdp(orc, golden):
    """The reference's CPU FC STDP reads the output spike through a pointer
    (`if (output(i, winner.x))`, src/spyker/cpu/stdp.cpp:60-66: Vec2's (i, j)
    returns T*), which is never null, so its output spike time is always 0.
    Its CUDA backend (src/spyker/cuda/stdp.cu:54-67) and the conv STDP read
    the value.  The oracle (and the device backend) follow the value reading;
    with every output neuron spiking at t=0 both readings coincide, which is
    what pins the rest of the update against the reference vector."""
    g = golden
    ones = np.ones_like(g["fcstdp_out"])
    k = orc.fcstdp(g["fcstdp_in"], g["fcstdp_kernel"], ones, [(0.004, -0.003)], g["fcstdp_win"], g["fcstdp_cnt"])
    assert np.array_equal(k, g["fcstdp_kernel_after"])
    # with the real output spikes, depression happens only where an input spikes
    # later than the neuron's first output spike
    k2 = orc.fcstdp(g["fcstdp_in"], g["fcstdp_kernel"], g["fcstdp_out"], [(0.004, -0.003)], g["fcstdp_win"],
                    g["fcstdp_cnt"])
    assert not np.array_equal(k2, k)
    assert k2.min() >= 0 and k2.max() <= 1


def _exact_update(w, pot, cfg):
    """The F64 formula evaluated with the roundings it has: (old-l)*(u-old)
    rounded to double, then one fused multiply-add rounded to double, then
    the clamp and the rounding to float."""
    pos, neg, stab, lo, hi = cfg
    old = Fraction(float(w))
    v = Fraction(float(Fraction(float(old - Fraction(lo))) * Fraction(float(Fraction(hi) - old)))) if stab else 1
    r = float(old + Fraction(pos if pot else neg) * v)  # fma: one rounding
    r = min(max(r, lo), hi) if r == r else r
    return np.float32(r)


def test_stdp_update_exact_rational(orc):
    rng = np.random.default_rng(3)
    cfgs = [(0.004, -0.003, True, 0.0, 1.0), (-0.002, 0.005, False, 0.1, 0.9), (0.15, -0.1125, True, 0.0, 1.0)]
    ws = np.concatenate([rng.random(300).astype(np.float32), np.float32([0, 1, 1e-30, 0.5, 0.9999999])])
    for cfg in cfgs:
        for w in ws:
            for pot in (0, 1):
                assert orc.stdp_update(float(w), pot, cfg) == _exact_update(w, pot, cfg)


def test_conv_against_float64_and_dyadic(orc):
    rng = np.random.default_rng(4)
    x = (rng.random((3, 4, 11, 13)) < 0.4).astype(np.uint8)
    k = rng.integers(0, 64, (5, 4, 3, 5)).astype(np.float32) / 64  # dyadic: every sum is exact in FP32
    for stride, pad in [(1, (0, 0, 0, 0)), (1, (1, 2, 1, 2)), (2, (2, 1, 0, 3))]:
        pt, pl, pb, pr = pad
        tx = torch.nn.functional.pad(torch.from_numpy(x).double(), (pl, pr, pt, pb))
        ref = torch.nn.functional.conv2d(tx, torch.from_numpy(k).double(), stride=stride).numpy()
        assert np.array_equal(orc.conv(x, k, stride, pad).astype(np.float64), ref)
        assert np.array_equal(orc.conv(x, k, stride, pad, f64=True), ref)
    xf = rng.random((2, 3, 9, 9)).astype(np.float32)
    kf = rng.random((4, 3, 3, 3)).astype(np.float32)
    ref = torch.nn.functional.conv2d(torch.from_numpy(xf).double(), torch.from_numpy(kf).double(), padding=1).numpy()
    np.testing.assert_allclose(orc.conv(xf, kf, 1, (1, 1, 1, 1)), ref, rtol=1e-5)
    np.testing.assert_allclose(orc.conv(xf, kf, 1, (1, 1, 1, 1), f64=True), ref, rtol=1e-13)


def test_fc_against_float64(orc):
    rng = np.random.default_rng(5)
    x = (rng.random((2, 3, 40)) < 0.3).astype(np.uint8)
    k = rng.integers(0, 256, (7, 40)).astype(np.float32) / 256
    assert np.array_equal(orc.fc(x, k).astype(np.float64), x.astype(np.float64) @ k.T.astype(np.float64))


def test_dog_gabor_against_float64(orc, golden):
    rng = np.random.default_rng(6)
    img = rng.random((2, 1, 20, 20)).astype(np.float32)
    g1, g2 = golden["gauss_7_1"], golden["gauss_7_2"]
    ker = np.stack([g1[None], g2[None]]).astype(np.float32)  # [2, 1, 7, 7]
    got = orc.dog(img, ker, (3, 3, 3, 3))
    tk = torch.from_numpy(ker.reshape(2, 1, 7, 7)).double()
    mid = torch.nn.functional.conv2d(torch.from_numpy(img).double(), tk, padding=3).numpy()
    np.testing.assert_allclose(got[:, 0], np.maximum(mid[:, 0] - mid[:, 1], 0), atol=2e-6)
    gk = np.stack([golden["gabor_7_0"], golden["gabor_7_45"]]).astype(np.float32)
    got = orc.gabor(img, gk, (3, 3, 3, 3))
    ref = torch.nn.functional.conv2d(torch.from_numpy(img).double(), torch.from_numpy(gk[:, None]).double(),
                                     padding=3).numpy()
    np.testing.assert_allclose(got, ref, atol=2e-5)


def test_log_against_float64(orc, golden):
    """LoG (cpu/dog.cpp:62-84): on - off is the Gaussian difference of an FP64
    conv, on and off are non-negative and never both non-zero, and the two
    halves land in [plane][2][count] order."""
    rng = np.random.default_rng(8)
    img = rng.random((3, 2, 19, 23)).astype(np.float32)
    g1, g2 = golden["gauss_7_1"], golden["gauss_7_2"]
    ker = np.stack([np.stack([g1, g2]), np.stack([g2, g1])]).astype(np.float32)  # [2, 2, 7, 7]
    got = orc.log(img, ker, (3, 2, 1, 3))
    assert got.shape == (3, 2 * 2 * 2, 17, 22)
    mid = torch.nn.functional.conv2d(torch.nn.functional.pad(torch.from_numpy(img).double().reshape(6, 1, 19, 23),
                                                              (2, 3, 3, 1)),
                                     torch.from_numpy(ker.reshape(4, 1, 7, 7)).double()).numpy()
    d = (mid[:, :2] - mid[:, 2:]).reshape(3, 2, 2, 17, 22)
    g = got.reshape(3, 2, 2, 2, 17, 22)
    on, off = g[:, :, 0], g[:, :, 1]
    assert (on >= 0).all() and (off >= 0).all() and not (on * off).any()
    np.testing.assert_allclose(on - off, d, atol=2e-6)
    # filter 1 is filter 0 with the Gaussians swapped: its on is filter 0's off
    np.testing.assert_allclose(on[:, :, 1], off[:, :, 0], atol=2e-6)
    # DoG is LoG's on half
    assert np.array_equal(orc.dog(img, ker, (3, 2, 1, 3)).reshape(3, 2, 2, 17, 22), on)


def test_zca_fit_closed_forms(orc):
    """ZCA fit (cpu/zca.cpp:66-97) against what the mathematics fixes."""
    # one feature: mean, the (rows - 1) divisor and eps as F32
    m, t = orc.zca_fit(np.array([[1.0], [2.0], [3.0], [4.0]]), 0.1)
    assert m[0] == 2.5
    np.testing.assert_allclose(t[0, 0], 1 / np.sqrt(5 / 3 + float(np.float32(0.1))), rtol=1e-14)
    assert abs(t[0, 0] - 1 / np.sqrt(5 / 3 + 0.1)) > 1e-11  # F32 eps, not the double
    # correlated features: trans is the unique symmetric positive-definite
    # matrix whose square inverts the textbook covariance + eps I
    rng = np.random.default_rng(10)
    x = rng.normal(size=(500, 6)) @ rng.normal(size=(6, 6)) + rng.normal(size=6)
    eps = 1e-3
    m, t = orc.zca_fit(x, eps)
    np.testing.assert_allclose(m, x.mean(axis=0), rtol=1e-13, atol=1e-13)
    cov = np.cov(x, rowvar=False, ddof=1) + float(np.float32(eps)) * np.eye(6)
    assert np.allclose(t, t.T, rtol=0, atol=1e-12)
    assert np.linalg.eigvalsh(t).min() > 0
    np.testing.assert_allclose(t @ t @ cov, np.eye(6), atol=1e-9)
    # the transform whitens: the covariance of (x - mean) trans is cov0 (cov0 + eps I)^-1
    y = orc.zca_trans(x, m, t)
    cov0 = np.cov(x, rowvar=False, ddof=1)
    np.testing.assert_allclose(np.cov(y, rowvar=False, ddof=1), cov0 @ np.linalg.inv(cov), atol=1e-9)
    # a constant feature: its eigenvalue (0, or a rounding-level negative) clamps to 0 -> 1/sqrt(eps)
    xc = x.copy()
    xc[:, 2] = 7.0
    m, t = orc.zca_fit(xc, 0.25)
    assert m[2] == 7.0
    np.testing.assert_allclose(t[2, 2], 2.0, rtol=1e-12)
    np.testing.assert_allclose(np.delete(t[2], 2), 0, atol=1e-12)


def test_zca_trans_reads_the_upper_triangle(orc):
    rng = np.random.default_rng(11)
    x = rng.normal(size=(40, 5))
    a = rng.normal(size=(5, 5))
    sym = a + a.T
    mean = rng.normal(size=5)
    np.testing.assert_allclose(orc.zca_trans(x, mean, sym), (x - mean) @ sym, rtol=1e-13, atol=1e-13)
    garbage = sym.copy()
    garbage[np.tril_indices(5, -1)] = 99.0
    assert np.array_equal(orc.zca_trans(x, mean, garbage), orc.zca_trans(x, mean, sym))


def test_zca_split_closed_forms(orc):
    rng = np.random.default_rng(12)
    x = rng.normal(size=(3, 4, 5)).astype(np.float32)
    s = orc.zca_split(x)
    assert s.shape == (3, 2, 4, 5)
    assert np.array_equal(s[:, 0] - s[:, 1], x) and not (s[:, 0] * s[:, 1]).any()
    assert (s >= 0).all()
    # unsigned -x wraps (C++: T(-x)), signed -(-128) stays -128 in int8
    assert orc.zca_split(np.array([[3, 0, 255]], np.uint8)).tolist() == [[[3, 0, 255], [253, 0, 1]]]
    assert orc.zca_split(np.array([[-128, 5, -7]], np.int8)).tolist() == [[[0, 5, 0], [0, 0, 7]]]


def test_mt19937_stream(orc):
    """The oracle's MT19937 against the standard's known answer (the 10000th output of
    std::mt19937 default-seeded with 5489 is 4123659995) and numpy's independent MT19937
    under the same seeding."""
    assert orc.mt_stream(5489, 10000)[-1] == 4123659995
    for seed in (0, 1, 12345, 2**32 - 1):
        bg = np.random.MT19937(0)
        bg._legacy_seeding(seed)
        assert np.array_equal(orc.mt_stream(seed, 2000), bg.random_raw(2000).astype(np.uint32))


def test_rate_code_against_reference_vectors(orc, golden):
    """The Poisson rate code (cpu/code.cpp:168-230) on the reference's own run with the same
    table, both time orders; the seed base wraps past 2^32 inside the batch."""
    x, table = golden["rate_in"], golden["rate_table"]
    for sort, key in ((False, "rate_code_latency_T15"), (True, "rate_code_sort_T15")):
        got = orc.rate_code(x, 15, sort, 2**32 - 2, table)
        assert np.array_equal(got, golden[key]), key
        steps = np.diff(got.astype(np.int32), axis=1)
        assert ((steps == 0) | (steps == 1)).all()  # counts grow by at most one spike a step
    # a table with too few rows is refused rather than read past
    with pytest.raises(RuntimeError):
        orc.rate_code(x, 15, True, 7, table[:16])


def test_rate_fire_against_reference_vectors(orc, golden):
    x = golden["rate_fire_in"]
    for th, key in ((0.1, "rate_fire_th01"), (0.0, "rate_fire_th0")):
        assert np.array_equal(orc.rate_fire(x, th), golden[key]), key
    # closed form on a ramp: x_t = t * 0.25 fires every step above the threshold gap
    ramp = (np.arange(8, dtype=np.float32) * 0.25).reshape(1, 8, 1)
    assert orc.rate_fire(ramp, 0.1).ravel().tolist() == [0, 1, 2, 3, 4, 5, 6, 7]
    assert orc.rate_fire(ramp, 0.3).ravel().tolist() == [0, 0, 1, 1, 2, 2, 3, 3]


def test_host_kernels(orc, golden):
    for kk, sd in [(3, 1.0), (3, 2.0), (7, 1.0), (7, 2.0)]:
        g = orc.gaussian_kernel(kk, sd)
        assert np.array_equal(g, golden["gauss_%d_%g" % (kk, sd)])
        assert abs(g.sum() - 1) < 1e-12
    import math
    for th in (0, 45, 90, 135):
        for kk, lam in [(7, 4.0), (11, 8.0)]:
            g = orc.gabor_kernel(kk, 0.56 * lam, math.radians(th), 0.5, lam, 0.0)
            assert np.array_equal(g, golden["gabor_%d_%d" % (kk, th)])


def test_oracle_vs_live_reference(orc, ref):
    """Where the reference build is present, random inputs agree too."""
    rng = np.random.default_rng(7)
    x = rng.random((4, 333)).astype(np.float32)
    x[x < 0.5] = 0
    assert np.array_equal(orc.rank_code(x, 15), ref.rank_code(x, 15))
    p = np.floor(rng.random((2, 5, 9, 6, 6)) * 30).astype(np.float32)
    th = orc.threshold(p, 14.0)
    assert np.array_equal(orc.inhibit(th), ref.inhibit(th))
    a, b = orc.convwta(orc.inhibit(th), 1, 4), ref.convwta(ref.inhibit(th), 1, 4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_readout_gather_scatter(orc, golden):
    """rank/rate gather and rank scatter (cpu/time.cpp): reference vectors, and the
    round trips that fix them: scatter(t) is the cumulative spike train of t, and
    gather of a cumulative {0,1} train is its first spike."""
    rd, rf = golden["readout_in_u8"], golden["readout_in_f32"]
    for th in (0, 2):
        assert np.array_equal(orc.rank_gather(rd, th), golden["rank_gather_u8_th%d" % th])
    assert np.array_equal(orc.rank_gather(rf, 0.5), golden["rank_gather_f32_th05"])
    assert np.array_equal(orc.rate_gather(rf), golden["rate_gather_f32"])
    tu, tf = golden["scatter_in_u8"], golden["scatter_in_f32"]
    assert np.array_equal(orc.rank_scatter(tu, 10), golden["rank_scatter_u8_T10"])
    assert np.array_equal(orc.rank_scatter(tf, 10), golden["rank_scatter_f32_T10"])
    # closed forms
    first = np.where((rd > 2).any(axis=1), (rd > 2).argmax(axis=1), rd.shape[1])
    assert np.array_equal(orc.rank_gather(rd, 2), first)
    steps = np.arange(10).reshape(1, 10, 1)
    assert np.array_equal(orc.rank_scatter(tf, 10), (steps >= tf[:, None, :]).astype(np.uint8))
    # round trips
    cum = orc.rank_scatter(tu, 10)
    assert np.array_equal(orc.rank_gather(cum), np.minimum(tu, 10))
    assert np.array_equal(orc.rank_scatter(orc.rank_gather(cum).astype(np.uint8), 10), cum)
    assert np.array_equal(orc.rate_gather(rf), rf[:, -1])
