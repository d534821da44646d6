"""spk_stdp_layer_fused_noalloc against spk_stdp_layer_fused, through the C ABI.

The allocation-free entry point carves its buffers from the library's
device scratch array instead of the stream's workspace; it must give the
allocating entry point's spikes, pooled spikes, winners, counts and
STDP-updated weights bit for bit, on two calls in a row (the second with a
smaller batch, so the array holds the first call's leftovers), and refuse
with an error what it cannot run without allocating.
"""

import numpy as np
import pytest
import torch

import spyker_amd as spk

pytestmark = pytest.mark.gpu


def step(name, x, kern, th, radius, count, pool, pad):
    """One fused step through the named entry point on copies of the weights."""
    n, t, cin, h, w = x.shape
    cout, _, kh, kw = kern.shape
    pt, pl, pb, pr = pad
    ho, wo = h + pt + pb - kh + 1, w + pl + pr - kw + 1
    pho, pwo = (ho - pool) // pool + 1, (wo - pool) // pool + 1
    k = kern.clone()
    dev = x.device
    out = torch.empty((n, t, cout, ho, wo), dtype=torch.uint8, device=dev)
    pooled = torch.empty((n, t, cout, pho, pwo), dtype=torch.uint8, device=dev)
    data = torch.empty((n, count, 5), dtype=torch.int64, device=dev)
    cnt = torch.empty((n,), dtype=torch.int64, device=dev)
    cfg, ncfg = spk._configs([spk.STDPConfig(0.004, -0.003)])
    spk.call(name, spk._ptr(x), n, t, cin, h, w, spk._ptr(k), cout, kh, kw, pt, pl, pb, pr, float(th), radius, radius,
             count, cfg, ncfg, spk._ptr(out), spk._ptr(pooled), pool, pool, pool, pool, spk._ptr(data), spk._ptr(cnt),
             None, spk._stream(x))
    torch.cuda.synchronize()
    return out, pooled, data, cnt, k


def spikes(rng, shape, p=0.5):
    s = (rng.random(shape) < p / shape[1] * 3).astype(np.uint8)
    return np.maximum.accumulate(s, axis=1)


def test_noalloc_matches_the_allocating_entry_point(dev):
    rng = np.random.default_rng(71)
    kern = torch.from_numpy(np.clip(rng.normal(0.8, 0.05, (250, 30, 3, 3)), 0, 1).astype(np.float32)).to(dev)
    # the layer-2 shape, then a smaller batch of another shape on the same array
    for n, t, cin, h in [(64, 15, 30, 14), (9, 12, 30, 11)]:
        x = torch.from_numpy(spikes(rng, (n, t, cin, h, h))).to(dev)
        want = step("spk_stdp_layer_fused", x, kern, 10.0, 1, 8, 3, (1, 1, 1, 1))
        got = step("spk_stdp_layer_fused_noalloc", x, kern, 10.0, 1, 8, 3, (1, 1, 1, 1))
        assert want[3].sum().item() > 0
        for a, b in zip(got, want):
            assert torch.equal(a, b)


def test_noalloc_refuses_what_needs_allocation(dev):
    rng = np.random.default_rng(72)
    kern = torch.from_numpy(np.clip(rng.normal(0.8, 0.05, (250, 30, 3, 3)), 0, 1).astype(np.float32)).to(dev)
    # buffers beyond the 16 MiB array: refused before any launch
    x = torch.zeros((2048, 15, 30, 14, 14), dtype=torch.uint8, device=dev)
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        step("spk_stdp_layer_fused_noalloc", x, kern, 10.0, 1, 8, 3, (1, 1, 1, 1))
    # a sample too large for the LDS would take the region path, whose plan allocates
    k7 = torch.from_numpy(np.clip(rng.normal(0.8, 0.05, (64, 8, 7, 7)), 0, 1).astype(np.float32)).to(dev)
    x = torch.zeros((1, 12, 8, 48, 48), dtype=torch.uint8, device=dev)
    with pytest.raises(RuntimeError, match="region path"):
        step("spk_stdp_layer_fused_noalloc", x, k7, 10.0, 5, 10, 2, (3, 3, 3, 3))
    # the entry point still works after the refusals
    x = torch.from_numpy(spikes(rng, (4, 15, 30, 14, 14))).to(dev)
    want = step("spk_stdp_layer_fused", x, kern, 10.0, 1, 8, 3, (1, 1, 1, 1))
    got = step("spk_stdp_layer_fused_noalloc", x, kern, 10.0, 1, 8, 3, (1, 1, 1, 1))
    for a, b in zip(got, want):
        assert torch.equal(a, b)
