"""Rate coding on the device: the Poisson rate code and rate fire.

Bars (bit for bit):
  * spk.code(..., code="rate") equals the oracle's restatement of the
    reference's CPU rate code (cpu/code.cpp:134-243, itself pinned to the
    reference's own run in test_oracle.py) on the device's table
    (spk_poisson_table with T + 1 rows; spike times past T read row T), with
    rnd the generator's draw for the call: the first draw after
    spk.random_seed(seed) for the first call, the second for the next;
  * spk.fire(..., code="rate") equals the oracle's rate fire
    (cpu/fire.cpp:84-114) and the reference vectors.
"""

import ctypes as C

import numpy as np
import pytest
import torch

import spyker_amd as spk
from spyker_amd import _lib

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


def draws(seed, k):
    """The first k outputs of std::mt19937(seed): the shared generator after spk.random_seed."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return [int(v) for v in bg.random_raw(k)]


def table_for(x, time, sort):
    """The device table (time + 1 rows) with row `time` repeated for every spike time past it:
    latency times never pass `time`; sorted times reach (len - 1) * (time + 1e-5) / #above-minimum."""
    t = np.zeros((time + 1, 1024), np.uint16)
    assert _lib.load().spk_poisson_table(t.ctypes.data_as(C.c_void_p), C.c_int64(time + 1), C.c_int64(1024)) == 0
    rows = time + 1
    if sort:
        above = min(int((row > row.min()).sum()) for row in x)
        rows = max(rows, min(65536, int((x.shape[1] - 1) * (time + 1) / max(1, above)) + 2))
    return np.concatenate([t, np.repeat(t[-1:], rows - (time + 1), axis=0)])


def inputs(rng, n, length):
    x = rng.random((n, length)).astype(np.float32)
    x[x < 0.3] = 0  # a block of ties at the minimum
    return x


@pytest.mark.parametrize("n,length,time", [(8, 300, 15), (3, 2000, 30), (2, 9000, 20), (5, 1, 4)])
@pytest.mark.parametrize("sort", [False, True])
def test_rate_code_matches_oracle(dev, orc, n, length, time, sort):
    rng = np.random.default_rng(n * 1000 + length + time)
    x = inputs(rng, n, length)
    if length > 1 and not sort:
        x[0] = 0.25  # a constant sample (latency scale is infinite)
    table = table_for(x, time, sort)
    seed = 1234 + length
    spk.random_seed(seed)
    rnd = draws(seed, 2)
    got1 = H(spk.code(T(x, dev), time, sort, code="rate"))
    got2 = H(spk.code(T(x, dev), time, sort, dtype="f32", code="RATE"))
    ref1 = orc.rate_code(x, time, sort, rnd[0], table)
    ref2 = orc.rate_code(x, time, sort, rnd[1], table)
    assert got1.shape == (n, time, length)
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got2, ref2.astype(np.float32))
    if length > 1:
        assert not np.array_equal(ref1, ref2)  # the second call draws anew
        assert ref1[:, -1].sum() > 0


def test_rate_code_sample_shapes(dev, orc):
    """[N, C, H, W] input codes its flattened samples; the table for a T is built once."""
    rng = np.random.default_rng(5)
    x = inputs(rng, 6, 2 * 28 * 28).reshape(6, 2, 28, 28)
    spk.random_seed(99)
    rnd = draws(99, 1)[0]
    got = H(spk.code(T(x, dev), 15, True, code="rate"))
    assert got.shape == (6, 15, 2, 28, 28)
    ref = orc.rate_code(x.reshape(6, -1), 15, True, rnd, table_for(x.reshape(6, -1), 15, True))
    assert np.array_equal(got.reshape(6, 15, -1), ref)
    with pytest.raises(RuntimeError):
        spk.code(T(x, dev), 15, code="poisson")


def test_rate_fire(dev, orc, golden):
    x = golden["rate_fire_in"]
    for th, key in ((0.1, "rate_fire_th01"), (0.0, "rate_fire_th0")):
        got = H(spk.fire(T(x, dev), th, code="rate"))
        assert np.array_equal(got, golden[key]), key
        assert np.array_equal(H(spk.fire(T(x, dev), th, dtype="f32", code="rate")), golden[key].astype(np.float32))
    rng = np.random.default_rng(6)
    x = np.cumsum(rng.random((4, 20, 3, 5, 5)).astype(np.float32), axis=1)
    assert np.array_equal(H(spk.fire(T(x, dev), 0.7, code="rate")), orc.rate_fire(x, 0.7))
    # rows of a multiple of 4 elements take the 4-wide path (F32 in, U8 and F32 out)
    x4 = np.cumsum(rng.random((5, 17, 2, 64)).astype(np.float32), axis=1)
    ref4 = orc.rate_fire(x4, 0.6)
    assert np.array_equal(H(spk.fire(T(x4, dev), 0.6, code="rate")), ref4)
    assert np.array_equal(H(spk.fire(T(x4, dev), 0.6, dtype="f32", code="rate")), ref4.astype(np.float32))
    # integer input: value + threshold < x in integer arithmetic (threshold converted to U8)
    u = np.cumsum(rng.integers(0, 3, (3, 16, 40)), axis=1).clip(0, 255).astype(np.uint8)
    assert np.array_equal(H(spk.fire(T(u, dev), 2, code="rate")), orc.rate_fire(u.astype(np.float32), 2.0))
