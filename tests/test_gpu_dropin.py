"""The reference's own layer code on the HIP backend (the drop-in check).

oracle/_ref/libspyker_dropin.so holds the reference's public C++ API and
dispatch (opers.cpp, impl.cpp, utils.cpp, shape.cpp), compiled from its
sources, with its Core::cuda_* seam implemented by
spyker_amd/seam/hip_backend.cpp over libspyker_hip.so.  These tests drive the
Kheradpisheh layer as the reference's examples do (DoG -> threshold -> code ->
Conv -> threshold -> inhibit -> convwta -> fire -> Conv.stdp -> pool) on
torch-ROCm tensors wrapped by address, through that unchanged reference
code, and compare every stage with the oracle bit for bit.
"""

import numpy as np
import pytest
import torch

import dropin
from spyker_amd import synthetic

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not dropin.available(), reason="oracle/_ref/libspyker_dropin.so is not built")]


@pytest.fixture(scope="module")
def ref_api():
    return dropin.DropIn()


def H(t):
    return t.cpu().numpy()


def test_reference_api_sees_the_device(dev, ref_api):
    assert ref_api.device_count() >= 1


def test_kheradpisheh_layer_through_the_reference_api(dev, orc, ref_api):
    n, t, size = 8, 15, 28
    img = (synthetic.strokes(n, size, seed=21).astype(np.float32) / 255).reshape(n, 1, size, size)
    x = torch.from_numpy(img).to(dev)

    # DoG on/off pairs at 3x3 and 7x7, same pads (the reference's DoG module on the device)
    f3 = torch.zeros((n, 2, size, size), dtype=torch.float32, device=dev)
    f7 = torch.zeros((n, 2, size, size), dtype=torch.float32, device=dev)
    ref_api.dog(0, 1, [(1.0, 2.0), (2.0, 1.0)], (1, 1, 1, 1), x, f3)
    ref_api.dog(0, 3, [(1.0, 2.0), (2.0, 1.0)], (3, 3, 3, 3), x, f7)
    g3 = np.stack([orc.gaussian_kernel(3, s) for s in (1.0, 2.0)] + [orc.gaussian_kernel(3, s) for s in (2.0, 1.0)])
    g7 = np.stack([orc.gaussian_kernel(7, s) for s in (1.0, 2.0)] + [orc.gaussian_kernel(7, s) for s in (2.0, 1.0)])
    # kernel bank [2][count][k][k]: first Gaussians (std1 of each filter), then second (std2)
    k3 = np.stack([g3[[0, 2]], g3[[1, 3]]]).astype(np.float32)
    k7 = np.stack([g7[[0, 2]], g7[[1, 3]]]).astype(np.float32)
    of = np.concatenate([orc.dog(img, k3, (1,) * 4), orc.dog(img, k7, (3,) * 4)], axis=1)
    feat = torch.cat([f3, f7], dim=1).contiguous()
    assert np.array_equal(H(feat), of)

    # threshold 0.01, rank code T=15 (u8 spikes)
    ref_api.threshold(feat, 0.01)
    of = orc.threshold(of, 0.01)
    spikes = torch.zeros((n, t, 4, size, size), dtype=torch.uint8, device=dev)
    ref_api.code(feat, spikes, t)
    ospk = orc.rank_code(of.reshape(n, -1), t).reshape(n, t, 4, size, size)
    assert np.array_equal(H(spikes), ospk)

    # Conv 4->30 5x5 pad 2 with the caller's weights (the module's kernel is the torch tensor)
    rng = np.random.default_rng(22)
    kern = np.clip(rng.normal(0.8, 0.05, (30, 4, 5, 5)), 0, 1).astype(np.float32)
    kt = torch.from_numpy(kern).to(dev)
    conv = dropin.Conv(ref_api, 0, 4, 30, 5, 1, (2, 2, 2, 2))
    conv.set_kernel(kt)
    conv.add_config(0.004, -0.003)
    pot = torch.zeros((n, t, 30, size, size), dtype=torch.float32, device=dev)
    conv.forward(spikes, pot)  # the reference casts the u8 spikes to the kernel's f32 first
    opot = orc.conv(ospk.reshape(n * t, 4, size, size), kern, 1, (2, 2, 2, 2)).reshape(n, t, 30, size, size)
    assert np.array_equal(H(pot), opot)

    # threshold 15, inhibit, convwta r3 k5, fire, STDP, pool 2x2
    ref_api.threshold(pot, 15.0)
    ref_api.inhibit(pot)
    opot = orc.inhibit(orc.threshold(opot, 15.0))
    assert np.array_equal(H(pot), opot)
    win, cnt = ref_api.convwta(pot, 3, 5)
    owin, ocnt = orc.convwta(opot, 3, 5)
    assert np.array_equal(cnt, ocnt) and ocnt.sum() > 0
    for i in range(n):
        assert np.array_equal(win[i, :cnt[i]], owin[i, :ocnt[i]])
    out = torch.zeros((n, t, 30, size, size), dtype=torch.uint8, device=dev)
    ref_api.fire(pot, out)
    oout = orc.fire(opot)
    assert np.array_equal(H(out), oout)
    conv.stdp(spikes, win, cnt, out)
    ok = orc.convstdp(ospk, kern, oout, [(0.004, -0.003)], owin, ocnt, 1, (2, 2, 2, 2))
    assert np.array_equal(H(kt), ok)  # updated in place in the caller's tensor
    pooled = torch.zeros((n, t, 30, 14, 14), dtype=torch.uint8, device=dev)
    ref_api.pool(out, pooled, 2, 2)
    assert np.array_equal(H(pooled), orc.pool(oout, 2, 2))


def test_reference_checks_and_errors_pass_through(dev, ref_api):
    """The reference's own shape checks run above the seam, and a failure in
    the backend surfaces in the reference's error form."""
    x = torch.zeros((2, 3, 4, 10, 10), dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="Input and output must have the same data type"):
        ref_api.pool(x, torch.zeros((2, 3, 4, 3, 3), dtype=torch.uint8, device=dev), 3, 3)
    with pytest.raises(RuntimeError, match="Tensor shapes are not compatible"):
        ref_api.pool(x, torch.zeros((2, 3, 4, 4, 4), dtype=torch.float32, device=dev), 3, 3)
    # a device dtype the backend does not run: the library's message in SpykerAssert's form
    xi = torch.zeros((2, 3, 4, 10, 10), dtype=torch.int32, device=dev)
    with pytest.raises(RuntimeError, match="Failure At: .*Scope: Core::"):
        ref_api.inhibit(xi)


def test_pad_through_the_reference_api(dev, golden, ref_api):
    for kind, pad, value, key in [("u8", (2, 0, 3, 1), 7, "pad_u8_p2031_v7"),
                                  ("f32", (3, 2, 1, 0), -1.5, "pad_f32_p3210_vm1p5")]:
        x = torch.from_numpy(golden["pad_in_" + kind]).to(dev)
        want = golden[key]
        out = torch.zeros(want.shape, dtype=x.dtype, device=dev)
        ref_api.pad(x, out, pad, value)
        assert np.array_equal(H(out), want)
