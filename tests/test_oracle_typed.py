"""Pins the typed (F16 / F32 / F64) numpy restatement of the plain path
(oracle/typed.py): at F32 it equals the C oracle; at F16 and F64 it equals
the reference's own CPU code (oracle/_ref, compiled from its sources) where
that is built; the F64-kernel conv of the C oracle against closed forms."""

import numpy as np
import pytest

import oracle
from oracle import typed


@pytest.fixture(scope="module")
def ref_t():
    if not oracle.reference_available():
        pytest.skip("reference build (oracle/_ref) not present")
    return oracle.ReferenceTyped()


def potentials(rng, shape, dt):
    """Random potentials with ties, zeros and values near a threshold of 2."""
    x = rng.choice([0.0, 0.5, 1.25, 2.0, 2.0009765625, 3.0, 4.5, 2.2, 0.7], size=shape)
    x = x + rng.normal(0, 0.3, size=shape) * (rng.random(shape) < 0.5)
    return np.cumsum(x, axis=1).astype(dt) if len(shape) > 2 else x.astype(dt)


def spikes(rng, shape, p=0.3):
    return np.cumsum(rng.random(shape) < p, axis=1).clip(0, 1).astype(np.uint8)


def test_typed_at_f32_equals_the_c_oracle(orc):
    rng = np.random.default_rng(61)
    x = potentials(rng, (3, 6, 5, 7, 8), np.float32)
    assert np.array_equal(typed.inhibit(x, 2.0), orc.inhibit(x, 2.0))
    assert np.array_equal(typed.threshold(x, 2.0), orc.threshold(x, 2.0))
    assert np.array_equal(typed.fire(x, 2.0), orc.fire(x, 2.0))
    for r, k, th in [(1, 4, 0.0), ((2, 1), 6, 3.0)]:
        w1, c1 = typed.convwta(x, r, k, th)
        w2, c2 = orc.convwta(x, r, k, th)
        assert np.array_equal(c1, c2) and c1.sum() > 0
        for i in range(x.shape[0]):
            assert np.array_equal(w1[i, :c1[i]], w2[i, :c2[i]])
    f = x.reshape(3, 6, -1)
    w1, c1 = typed.fcwta(f, 3, 5, 1.0)
    w2, c2 = orc.fcwta(f, 3, 5, 1.0)
    assert np.array_equal(c1, c2)
    for i in range(3):
        assert np.array_equal(w1[i, :c1[i]], w2[i, :c2[i]])
    si, so = spikes(rng, (2, 5, 3, 9, 9)), spikes(rng, (2, 5, 4, 9, 9), 0.2)
    kern = rng.uniform(0, 1, (4, 3, 3, 3)).astype(np.float32)
    win = np.zeros((2, 3), oracle.WINNER)
    win[0, 0] = (0, 0, 1, 2, 3)
    win[0, 1] = (1, 0, 3, 0, 8)
    win[1, 0] = (0, 0, 1, 4, 4)
    cnt = np.array([2, 1])
    cfg = [(0.004, -0.003), (0.01, -0.02, False, 0.1, 0.9)]
    assert np.array_equal(typed.convstdp(si, kern, so, cfg, win, cnt, 1, (1, 1, 1, 1)),
                          orc.convstdp(si, kern, so, cfg, win, cnt, 1, (1, 1, 1, 1)))


@pytest.mark.parametrize("dt", [np.float16, np.float64])
def test_typed_equals_the_reference(ref_t, dt):
    rng = np.random.default_rng(62)
    x = potentials(rng, (3, 6, 5, 7, 8), dt)
    assert np.array_equal(typed.threshold(x, 2.0), ref_t.threshold(x, 2.0))
    assert np.array_equal(typed.fire(x, 2.0), ref_t.fire(x, 2.0))
    got, want = typed.inhibit(x, 2.0), ref_t.inhibit(x, 2.0)
    assert np.array_equal(got, want) and np.array_equal(np.signbit(got), np.signbit(want))
    for r, k, th in [(1, 4, 0.0), ((2, 1), 6, 3.0), (0, 40, 100.0)]:
        w1, c1 = typed.convwta(x, r, k, th)
        w2, c2 = ref_t.convwta(x, r, k, th)
        assert np.array_equal(c1, c2) and (th > 10 or c1.sum() > 0)
        for i in range(x.shape[0]):
            assert np.array_equal(w1[i, :c1[i]], w2[i, :c2[i]])
    f = x.reshape(3, 6, -1)
    w1, c1 = typed.fcwta(f, 3, 5, 1.0)
    w2, c2 = ref_t.fcwta(f, 3, 5, 1.0)
    assert np.array_equal(c1, c2)
    for i in range(3):
        assert np.array_equal(w1[i, :c1[i]], w2[i, :c2[i]])
    si, so = spikes(rng, (2, 5, 3, 9, 9)), spikes(rng, (2, 5, 4, 9, 9), 0.2)
    kern = rng.uniform(0, 1, (4, 3, 3, 3)).astype(dt)
    win = np.zeros((2, 3), oracle.WINNER)
    win[0, 0] = (0, 0, 1, 2, 3)
    win[0, 1] = (1, 0, 3, 0, 8)
    win[1, 0] = (0, 0, 1, 4, 4)
    cnt = np.array([2, 1])
    cfg = [(0.004, -0.003), (0.01, -0.02, False, 0.1, 0.9)]
    assert np.array_equal(typed.convstdp(si, kern, so, cfg, win, cnt, 1, (1, 1, 1, 1)),
                          ref_t.convstdp(si, kern, so, cfg, win, cnt, 1, (1, 1, 1, 1)))


def test_f16_wta_unfired_value_is_the_smallest_normal():
    """Limits<half>::min() is 2^-14, not the lowest half: a sample whose
    neurons never pass the threshold yields no winner, and a fired neuron
    whose value rounds to 0 still beats the unfired ones."""
    x = np.zeros((1, 3, 1, 2, 2), np.float16)
    w, c = typed.convwta(x, 0, 4, 1.0)
    assert c[0] == 0
    x[0, 2, 0, 0, 1] = 2.0  # fires at the last step, value (T-1-2) + (2-0)/(2*2) = 0.5
    w, c = typed.convwta(x, 0, 4, 1.0)
    assert c[0] == 1 and tuple(w[0, 0])[2:] == (0, 0, 1)


def test_conv_f64k_closed_forms(orc):
    rng = np.random.default_rng(63)
    # dyadic inputs and weights: every order of the sum is exact
    x = rng.integers(0, 4, (2, 3, 9, 10)).astype(np.float64) / 4
    k = rng.integers(-8, 9, (5, 3, 3, 3)).astype(np.float64) / 16
    got = orc.conv_f64k(x, k, 1, (1, 1, 1, 1))
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    want = np.zeros_like(got)
    for a in range(3):
        for b in range(3):
            want += np.einsum("ncyx,oc->noyx", xp[:, :, a:a + 9, b:b + 10], k[:, :, a, b])
    assert np.array_equal(got, want)
    # one tap: out = x * k rounded once
    x1 = rng.normal(size=(1, 1, 4, 4))
    k1 = rng.normal(size=(1, 1, 1, 1))
    assert np.array_equal(orc.conv_f64k(x1, k1), x1 * k1[0, 0, 0, 0])
    # two taps: the second is a fused multiply-add of the first product
    x2 = np.array([[[[1.0 + 2.0 ** -30, 1.0 - 2.0 ** -30]]]])
    k2 = np.array([[[[1.0 + 2.0 ** -30, -(1.0 - 2.0 ** -30)]]]])
    p1 = (1.0 + 2.0 ** -30) * (1.0 + 2.0 ** -30)  # rounded: drops 2^-60
    exact = (1 + 2.0 ** -30) ** 2 - (1 - 2.0 ** -30) ** 2  # 2^-28, what one fma of p1 keeps
    got = orc.conv_f64k(x2, k2)[0, 0, 0, 0]
    assert got == typed._fma(1.0 - 2.0 ** -30, -(1.0 - 2.0 ** -30), p1) and got != p1 - (1 - 2.0 ** -30) ** 2
    assert abs(got - exact) <= 2.0 ** -60
    # F32-representable data: equals the F32-kernel F64 chain (exact products)
    xf = rng.normal(size=(2, 2, 6, 6)).astype(np.float32)
    kf = rng.normal(size=(3, 2, 3, 3)).astype(np.float32)
    assert np.array_equal(orc.conv_f64k(xf, kf.astype(np.float64)), orc.conv(xf, kf, f64=True))


def test_conv_f16_rounds_the_f32_sum_once(orc):
    rng = np.random.default_rng(64)
    x = (rng.random((2, 3, 8, 8)) < 0.4).astype(np.uint8)
    k = (rng.integers(0, 64, (4, 3, 3, 3)) / 64).astype(np.float16)
    got = typed.conv(orc, x, k, 1, (1, 1, 1, 1))
    assert got.dtype == np.float16
    want = orc.conv(x, k.astype(np.float32), 1, (1, 1, 1, 1))  # exact: multiples of 1/64 below 2^11/64
    assert np.array_equal(got, want.astype(np.float16))
    assert np.array_equal(got.astype(np.float32), want)


@pytest.mark.parametrize("dt", [np.float16, np.float32, np.float64])
def test_typed_pool_pad(orc, dt):
    rng = np.random.default_rng(65)
    x = np.where(rng.random((3, 4, 9, 11)) < 0.3, 0.0, rng.normal(0, 1, (3, 4, 9, 11)))
    x = np.where(rng.random(x.shape) < 0.1, -0.0, x).astype(dt)
    for k, s, p in [(2, 2, (0, 0, 0, 0)), (3, 2, (1, 0, 2, 1))]:
        got = typed.pool(x, k, s, p)
        if dt == np.float32:
            want = orc.pool(x, k, s, p)
            assert np.array_equal(got, want) and np.array_equal(np.signbit(got), np.signbit(want))
        elif oracle.reference_available():
            want = oracle.ReferenceTyped().pool(x, k, s, p)
            assert np.array_equal(got, want) and np.array_equal(np.signbit(got), np.signbit(want))
    if dt == np.float16:
        assert typed.pool(np.full((1, 2, 2), -1.0, dt), 2)[0, 0, 0] == np.float16(2.0 ** -14)
    got = typed.pad(x, (2, 0, 1, 3), -1.5)
    if dt == np.float32:
        assert np.array_equal(got, orc.pad(x, (2, 0, 1, 3), -1.5))
    elif oracle.reference_available():
        assert np.array_equal(got, oracle.ReferenceTyped().pad(x, (2, 0, 1, 3), -1.5))
