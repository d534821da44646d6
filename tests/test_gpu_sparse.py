"""Parity of the device sparse event path with the oracle and the reference vectors.

Bar: bit-exact -- event lists (offsets and packed events, list order
included), first-spike maps, winners and the STDP-updated F32 kernels.
Every call goes through the C ABI via spyker_amd.sparse.  The reference
vectors (tests/golden/sparse_vectors.npz) come from the reference's own
sparse code; larger seeded cases compare with the oracle, which
tests/test_sparse_oracle.py pins to those vectors.
"""

import os

import numpy as np
import pytest
import torch

import spyker_amd as spk
from oracle import WINNER, SparseSpikes
from spyker_amd import sparse, synthetic

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sparse_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.cpu().numpy()


def sp(g, name):
    return SparseSpikes(g[name + "_off"], g[name + "_ev"], tuple(int(v) for v in g[name + "_shape"]))


def host(s):
    """SparseTensor -> the oracle's SparseSpikes."""
    return SparseSpikes(H(s.offsets()), H(s.events()), tuple(s.shape))


def winners_np(w):
    d, c = H(w.data), H(w.counts)
    out = np.zeros(d.shape[:2], WINNER)
    for i, f in enumerate("ctzyx"):
        out[f] = d[:, :, i]
    return out, c


def winners_dev(w, counts, dev):
    arr = np.stack([w[f] for f in "ctzyx"], axis=-1).astype(np.int64)
    return spk.Winners(T(arr, dev), T(np.asarray(counts, np.int64), dev))


def same_winners(w, c, gw, gc):
    assert np.array_equal(c, gc)
    for i in range(len(c)):
        assert np.array_equal(w[i, :c[i]], gw[i, :gc[i]])


def test_sparse_code(dev, g, orc):
    img = T(g["code_in"], dev)
    assert host(sparse.code(img, 12, True)) == sp(g, "code_sort")
    assert host(sparse.code(img, 12, False)) == sp(g, "code_latency")
    rng = np.random.default_rng(11)
    x = (rng.random((5, 3, 40, 37)) * (rng.random((5, 3, 40, 37)) < 0.4)).astype(np.float32)
    x[1, :, :8, :8] = 0.5  # ties
    for sort in (True, False):
        for t in (1, 15, 300):
            assert host(sparse.code(T(x, dev), t, sort)) == orc.sparse_code(x, t, sort), (sort, t)


def test_sparse_dense_round_trip(dev, g, orc):
    d = g["dense_in"]
    for th in (0, 1):
        assert host(spk.SparseTensor(T(d, dev), th)) == sp(g, "from_dense_%d" % th)
    s = spk.SparseTensor(T(d, dev))
    assert s.numel == len(g["from_dense_0_ev"])
    assert np.array_equal(H(s.dense()), g["to_dense"])
    for dt in ("u8", "u16", "i32", "f32"):
        assert np.array_equal(H(sparse.gather(s, dt)).astype(np.int64), g["gather"].astype(np.int64)), dt
    # F32 input: the event is at the first step whose value is > threshold
    rng = np.random.default_rng(5)
    x = rng.random((3, 9, 4, 33, 35)).astype(np.float32)
    assert host(spk.SparseTensor(T(x, dev), 0.9)) == orc.sparse_from_dense((x > np.float32(0.9)).astype(np.uint8))
    # several builder chunks per sample, every step populated
    big = (rng.random((2, 20, 6, 64, 64)) < 0.01).astype(np.uint8)
    s = spk.SparseTensor(T(big, dev))
    assert host(s) == orc.sparse_from_dense(big)
    assert np.array_equal(H(s.dense()), orc.sparse_to_dense(orc.sparse_from_dense(big)))
    assert host(s.copy()) == host(s)


def test_sparse_empty(dev, orc):
    z = np.zeros((2, 5, 3, 7, 9), np.uint8)
    s = spk.SparseTensor(T(z, dev))
    assert s.numel == 0 and host(s) == orc.sparse_from_dense(z)
    assert not H(s.dense()).any()
    assert (H(sparse.gather(s)) == 5).all()
    assert host(sparse.pool(s, 2)).ev.size == 0 and host(sparse.inhibit(s)).ev.size == 0
    k = np.full((4, 3, 3, 3), 0.5, np.float32)
    assert host(sparse.conv(s, T(k, dev), 1.0, pad=1)).ev.size == 0
    w = sparse.convwta(s, 1, 3)
    assert (H(w.counts) == 0).all()


def test_sparse_pad_pool_inhibit(dev, g, orc):
    s = spk.SparseTensor(T(g["dense_in"], dev))
    assert host(sparse.pad(s, (2, 1, 2, 1))) == sp(g, "pad")
    assert host(sparse.pool(s, 2, pad=(1, 0, 0, 0))) == sp(g, "pool2_pad10")
    assert host(sparse.pool(s, 3)) == sp(g, "pool3")
    assert host(sparse.inhibit(s)) == sp(g, "inhibit")
    rng = np.random.default_rng(6)
    big = (rng.random((3, 15, 20, 50, 44)) < 0.02).astype(np.uint8)
    o = orc.sparse_from_dense(big)
    s = spk.SparseTensor(T(big, dev))
    for k, pad in ((2, (1, 0, 0, 0)), (3, (0, 0, 0, 0)), (4, (1, 2, 3, 0))):
        assert host(sparse.pool(s, k, pad=pad)) == orc.sparse_pool(o, k, pad), (k, pad)
    assert host(sparse.inhibit(s)) == orc.sparse_inhibit(o)
    with pytest.raises(RuntimeError, match="Stride must be equal"):
        sparse.pool(s, 2, 1)


def test_sparse_convwta(dev, g, orc):
    s = spk.SparseTensor(T(g["dense_in"], dev))
    for name, r, k in (("convwta_r1_k5", 1, 5), ("convwta_r02_k3", (0, 2), 3)):
        w, c = winners_np(sparse.convwta(s, r, k))
        same_winners(w, c, g[name + "_win"], g[name + "_cnt"])
    rng = np.random.default_rng(7)
    big = (rng.random((4, 15, 150, 12, 12)) < 0.01).astype(np.uint8)
    o = orc.sparse_from_dense(big)
    s = spk.SparseTensor(T(big, dev))
    for r, k in ((1, 8), (0, 100), (2, 1), ((3, 0), 40)):
        w, c = winners_np(sparse.convwta(s, r, k))
        gw, gc = orc.sparse_convwta(o, r, k)
        same_winners(w, c, gw, gc)


def test_sparse_conv(dev, g, orc):
    s, k = spk.SparseTensor(T(g["dense_in"], dev)), g["conv_kernel"]
    assert host(sparse.conv(s, T(k, dev), 6.0, pad=(2, 2, 2, 2))) == sp(g, "conv")
    assert host(sparse.conv(sparse.pool(s, 2), T(k, dev), 4.0, pad=(2, 2, 2, 2))) == sp(g, "conv_pooled")
    kn = np.ascontiguousarray(k[:3] - 0.9)
    assert host(sparse.conv(s, T(kn, dev), -0.5, pad=(1, 0, 2, 3))) == sp(g, "conv_negative")
    with pytest.raises(RuntimeError, match="Stride must be 1"):
        sparse.conv(s, T(k, dev), 6.0, stride=2, pad=2)


def test_sparse_conv_large(dev, orc):
    """40 output channels (two channel chunks), 3x3 output tiles with ragged
    edges, more than 256 events per list (several compaction rounds)."""
    rng = np.random.default_rng(8)
    d = (rng.random((2, 10, 2, 40, 40)) < 0.15).astype(np.uint8)
    o = orc.sparse_from_dense(d)
    s = spk.SparseTensor(T(d, dev))
    k = np.clip(rng.normal(0.8, 0.05, (40, 2, 5, 5)), 0, 1).astype(np.float32)
    for th, pad in ((8.0, (2, 2, 2, 2)), (20.0, (0, 1, 3, 2))):
        got = sparse.conv(s, T(k, dev), th, pad=pad)
        assert host(got) == orc.sparse_conv(o, k, th, pad), (th, pad)
    k2 = np.clip(rng.normal(0.8, 0.05, (35, 40, 2, 2)), 0, 1).astype(np.float32)
    mid = orc.sparse_conv(o, k, 8.0, (2, 2, 2, 2))
    got = sparse.conv(sparse.conv(s, T(k, dev), 8.0, pad=2), T(k2, dev), 1.0, pad=1)
    assert host(got) == orc.sparse_conv(mid, k2, 1.0, (1, 1, 1, 1))


def test_sparse_stdp(dev, g, orc):
    s, k = spk.SparseTensor(T(g["dense_in"], dev)), g["conv_kernel"]
    cfg = [spk.STDPConfig(0.004, -0.003), spk.STDPConfig(0.01, -0.002, False, -1.0, 2.0)]
    kd, w = T(k, dev), g["stdp_win"]
    sparse.conv_stdp(s, kd, cfg, winners_dev(w, g["stdp_cnt"], dev), (2, 2, 2, 2))
    assert np.array_equal(H(kd), g["stdp_kernel"])
    rng = np.random.default_rng(9)
    big = (rng.random((3, 12, 4, 30, 30)) < 0.03).astype(np.uint8)
    o = orc.sparse_from_dense(big)
    kb = np.clip(rng.normal(0.5, 0.2, (6, 4, 5, 5)), 0, 1).astype(np.float32)
    wb = np.zeros((3, 9), WINNER)  # many winners per channel: long update chains
    wb["c"], wb["t"] = rng.integers(-1, 2, (3, 9)), rng.integers(0, 12, (3, 9))
    wb["z"], wb["y"], wb["x"] = rng.integers(0, 6, (3, 9)), rng.integers(0, 30, (3, 9)), rng.integers(0, 30, (3, 9))
    cb = np.array([9, 4, 7], np.int64)
    kd = T(kb, dev)
    sparse.conv_stdp(spk.SparseTensor(T(big, dev)), kd, cfg, winners_dev(wb, cb, dev), (2, 2, 2, 2))
    cfg_o = [(0.004, -0.003), (0.01, -0.002, False, -1.0, 2.0)]
    assert np.array_equal(H(kd), orc.sparse_stdp(o, kb, cfg_o, wb, cb, (2, 2, 2, 2)))
    bad = w.copy()
    bad["c"][1, 0] = 2
    kd = T(k, dev)
    with pytest.raises(RuntimeError, match="configuration index"):
        sparse.conv_stdp(s, kd, cfg, winners_dev(bad, g["stdp_cnt"], dev), (2, 2, 2, 2))
    assert np.array_equal(H(kd), k)  # nothing applied


def test_fcwta_u8(dev, g, orc):
    dense = orc.sparse_to_dense(sp(g, "conv")).reshape(3, 10, -1)
    for r, cnt in ((0, 16), (3, 5)):
        w, c = winners_np(spk.fcwta(T(dense, dev), r, cnt))
        same_winners(w, c, g["fcwta_u8_r%d_k%d_win" % (r, cnt)].reshape(3, cnt), g["fcwta_u8_r%d_k%d_cnt" % (r, cnt)])


def test_sparse_layer_pipeline(dev, orc):
    """DVS-like events -> sparse conv 2->64 5x5 pad 2 -> inhibit -> convwta ->
    sparse STDP, and pooled output; every stage equal to the oracle's."""
    ev = synthetic.events(4, 15, 48, seed=21, density=0.04)
    o = orc.sparse_from_dense(ev)
    s = spk.SparseTensor(T(ev, dev))
    assert host(s) == o
    rng = np.random.default_rng(22)
    k = np.clip(rng.normal(0.8, 0.05, (64, 2, 5, 5)), 0, 1).astype(np.float32)
    layer = spk.Conv(2, 64, 5, pad=2, device=dev)
    layer.kernel.copy_(T(k, dev))
    layer.stdpconfig = [spk.STDPConfig(0.004, -0.003)]
    out = spk.sparse.inhibit(layer(s, 8.0))
    oo = orc.sparse_inhibit(orc.sparse_conv(o, k, 8.0, (2, 2, 2, 2)))
    assert host(out) == oo and len(oo.ev) > 0
    win = sparse.convwta(out, 2, 5)
    gw, gc = orc.sparse_convwta(oo, 2, 5)
    w, c = winners_np(win)
    same_winners(w, c, gw, gc)
    layer.stdp(s, win)
    assert np.array_equal(H(layer.kernel), orc.sparse_stdp(o, k, [(0.004, -0.003)], gw, gc, (2, 2, 2, 2)))
    pooled = sparse.pool(out, 2, pad=1)
    assert host(pooled) == orc.sparse_pool(oo, 2, (1, 1, 1, 1))
    fw, fcnt = winners_np(spk.fcwta(out.dense().reshape(4, 15, -1), 0, 16))
    gw, gc = orc.fcwta(orc.sparse_to_dense(oo).reshape(4, 15, -1), 0, 16)
    same_winners(fw, fcnt, gw.reshape(4, 16), gc)


def check_fcwta(orc, o, w, c, radius, count):
    """Sparse fcwta = the dense integer fcwta on the densified tensor (its
    flat index, c = 0), with t = the winner's spike step."""
    n, t, ch, h, wd = o.shape
    gw, gc = orc.fcwta(orc.sparse_to_dense(o).reshape(n, t, -1), radius, count)
    gw = gw.reshape(n, count)
    first = orc.sparse_gather(o).reshape(n, -1)
    assert np.array_equal(c, gc)
    for i in range(n):
        got = w[i, :c[i]]
        assert np.array_equal((got["z"] * h + got["y"]) * wd + got["x"], gw[i, :gc[i]]["x"])
        assert (got["c"] == 0).all()
        assert np.array_equal(got["t"], first[i, gw[i, :gc[i]]["x"]])


def test_sparse_fcwta(dev, g, orc):
    """Sorted lists (conv output) and unsorted ones (pooled lists keep the
    input's order), radius 0 and wider, counts past the events."""
    s = spk.SparseTensor(T(g["dense_in"], dev))
    conv_out = sparse.conv(s, T(g["conv_kernel"], dev), 6.0, pad=(2, 2, 2, 2))
    pooled = sparse.pool(s, 2, pad=(1, 0, 0, 0))
    assert host(conv_out) == sp(g, "conv")
    for x in (conv_out, pooled):
        o = host(x)
        for r, k in ((0, 16), (3, 5), (40, 7), (0, 5000)):
            w, c = winners_np(sparse.fcwta(x, r, k))
            check_fcwta(orc, o, w, c, r, k)
    # golden dense fcwta on the densified conv output
    for r, cnt in ((0, 16), (3, 5)):
        w, c = winners_np(sparse.fcwta(conv_out, r, cnt))
        assert np.array_equal(c, g["fcwta_u8_r%d_k%d_cnt" % (r, cnt)])
        flat = (w["z"] * 24 + w["y"]) * 24 + w["x"]
        gw = g["fcwta_u8_r%d_k%d_win" % (r, cnt)].reshape(3, cnt)
        for i in range(3):
            assert np.array_equal(flat[i, :c[i]], gw[i, :c[i]]["x"])


def test_sparse_fcwta_stdp_pipeline(dev, orc):
    """events -> sparse conv -> sparse fcwta -> sparse STDP, the event-stream
    layer step, against the oracle's dense fcwta on the densified output."""
    ev = synthetic.events(3, 20, 64, seed=23, density=0.02)
    o = orc.sparse_from_dense(ev)
    rng = np.random.default_rng(24)
    k = np.clip(rng.normal(0.8, 0.05, (64, 2, 5, 5)), 0, 1).astype(np.float32)
    s = spk.SparseTensor(T(ev, dev))
    kd = T(k, dev)
    out = sparse.conv(s, kd, 8.0, pad=2)
    oo = orc.sparse_conv(o, k, 8.0, (2, 2, 2, 2))
    assert host(out) == oo
    win = sparse.fcwta(out, 0, 16)
    w, c = winners_np(win)
    check_fcwta(orc, oo, w, c, 0, 16)
    sparse.conv_stdp(s, kd, [spk.STDPConfig(0.004, -0.003)], win, 2)
    assert np.array_equal(H(kd), orc.sparse_stdp(o, k, [(0.004, -0.003)], w, c, (2, 2, 2, 2)))


def test_sparse_scans_past_one_grid(dev, orc):
    """Prefix scans longer than one capped launch grid (4,194,304 entries):
    2048 x 2100 single-element lists in the builder, and a 4.6M-event filter
    (pool 1x1 keeps every event)."""
    rng = np.random.default_rng(25)
    d = (rng.random((2048, 2100, 1, 1, 1)) < 0.002).astype(np.uint8)
    d[-1, -1] = 1
    s = spk.SparseTensor(T(d, dev))
    o = orc.sparse_from_dense(d)
    assert host(s) == o and s.numel == len(o.ev)
    full = spk.SparseTensor(torch.ones((1, 1, 80, 240, 240), dtype=torch.uint8, device=dev))
    assert full.numel == 80 * 240 * 240
    same = sparse.pool(full, 1)
    assert same.numel == full.numel
    assert torch.equal(same.offsets(), full.offsets()) and torch.equal(same.events(), full.events())
