"""The C++ interface (include/spyker_amd/spyker.h) driven by tests/cpp/api_test.

The driver uses the interface the way a C++ user of the reference would
(Spyker::DoG, code, Conv, threshold, inhibit, convwta, fire, Conv::stdp,
pool, plus the fused layerStep) and writes what it computed; the results are
checked against the oracle here.
"""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "api_test")
LIB = os.path.join(ROOT, "spyker_amd", "libspyker_cpp.so")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/api_test is not built (python -m spyker_amd.build)")
    return DRIVER


def test_cpp_host_side():
    r = subprocess.run([_driver(), "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "host ok" in r.stdout, r.stdout + r.stderr


def test_cpp_library_exports_the_interface():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for name in ["Spyker::Conv::stdp(", "Spyker::Conv::Conv(", "Spyker::DoG::DoG(", "Spyker::Gabor::Gabor(",
                 "Spyker::FC::stdp(", "Spyker::code(", "Spyker::fire(", "Spyker::threshold(", "Spyker::inhibit(",
                 "Spyker::convwta(", "Spyker::fcwta(", "Spyker::pool(", "Spyker::conv(", "Spyker::fc(",
                 "Spyker::pad(", "Spyker::layerStep(", "Spyker::Tensor::to(", "Spyker::gather(", "Spyker::scatter(",
                 "Spyker::SparseTensor::SparseTensor(", "Spyker::SparseTensor::dense(", "Spyker::Sparse::conv(",
                 "Spyker::Sparse::pool(", "Spyker::Sparse::inhibit(", "Spyker::Sparse::convwta(",
                 "Spyker::Sparse::gather(", "Spyker::Sparse::code(", "Spyker::LoG::LoG(", "Spyker::ZCA::fit(",
                 "Spyker::ZCA::split("]:
        assert name in out, name


def _winners(path):
    rows = []
    for line in open(path):
        v = [int(x) for x in line.split()]
        rows.append([tuple(v[1 + 5 * k:6 + 5 * k]) for k in range(v[0])])
    return rows


@pytest.mark.gpu
def test_cpp_layer_matches_oracle(tmp_path, orc):
    from spyker_amd import synthetic
    n, t, cin, h, w, cout, k, pad, radius, count, poolk = 4, 15, 8, 14, 14, 40, 3, 1, 1, 5, 2
    rng = np.random.default_rng(21)
    img = (synthetic.strokes(n, 28, seed=21).astype(np.float32) / 255)
    x = np.maximum.accumulate((rng.random((n, t, cin, h, w)) < 0.04).astype(np.uint8), axis=1)
    kern = (rng.integers(40, 64, (cout, cin, k, k)) / 64).astype(np.float32)  # exact in FP16, sums exact in FP32
    pot = orc.conv(x.reshape(n * t, cin, h, w), kern, 1, (pad,) * 4)
    th = float(np.quantile(pot, 0.9)) + 0.5 / 64
    d = str(tmp_path)
    img.tofile(os.path.join(d, "img.f32"))
    x.tofile(os.path.join(d, "x.u8"))
    kern.tofile(os.path.join(d, "k.f32"))
    with open(os.path.join(d, "shape.txt"), "w") as f:
        f.write("%d %d %d %d %d %d %d %d %.17g %d %d %d %.17g\n" % (n, t, cin, h, w, cout, k, pad, th, radius, count,
                                                                   poolk, 0.0))
    r = subprocess.run([_driver(), "layer", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "layer ok" in r.stdout, r.stdout + r.stderr

    def rd(name, dtype, shape):
        return np.fromfile(os.path.join(d, name), dtype=dtype).reshape(shape)

    # front end: DoG (same pad 1, sigma pairs (1,2), (2,1)), threshold 0.01, rank code
    g1, g2 = orc.gaussian_kernel(3, 1.0), orc.gaussian_kernel(3, 2.0)
    dk = np.stack([np.stack([g1, g2]), np.stack([g2, g1])]).astype(np.float32)
    dog = orc.threshold(orc.dog(img.reshape(n, 1, 28, 28), dk, (1, 1, 1, 1)), 0.01)
    assert np.array_equal(rd("dog.out", np.float32, dog.shape), dog)
    assert np.array_equal(rd("code.out", np.uint8, (n, t, 2, 28, 28)),
                          orc.rank_code(dog.reshape(n, -1), t).reshape(n, t, 2, 28, 28))

    # plain layer step
    ho = h + 2 * pad - k + 1
    opot = orc.inhibit(orc.threshold(pot.reshape(n, t, cout, ho, ho), th))
    ow, oc = orc.convwta(opot, radius, count)
    oout = orc.fire(opot)
    ok = orc.convstdp(x, kern, oout, [(0.004, -0.003)], ow, oc, 1, (pad,) * 4)
    assert oc.sum() > 0
    assert np.array_equal(rd("pot.out", np.float32, opot.shape), opot)
    assert np.array_equal(rd("out.out", np.uint8, oout.shape), oout)
    assert np.array_equal(rd("kernel.out", np.float32, kern.shape), ok)
    assert np.array_equal(rd("pooled.out", np.uint8, (n, t, cout, ho // poolk, ho // poolk)), orc.pool(oout, poolk))
    want = [[(0, 0, int(ow[i, j]["z"]), int(ow[i, j]["y"]), int(ow[i, j]["x"])) for j in range(oc[i])]
            for i in range(n)]
    assert _winners(os.path.join(d, "winners.out")) == want

    # the fused step from the same weights: exact on these dyadic weights
    assert np.array_equal(rd("fused_out.out", np.uint8, oout.shape), oout)
    assert np.array_equal(rd("fused_kernel.out", np.float32, kern.shape), ok)
    assert np.array_equal(rd("fused_pooled.out", np.uint8, (n, t, cout, ho // poolk, ho // poolk)),
                          orc.pool(oout, poolk))
    assert _winners(os.path.join(d, "fused_winners.out")) == want


@pytest.mark.gpu
def test_cpp_sparse_matches_oracle(tmp_path, orc):
    from spyker_amd import synthetic
    n, t, h, w, cout, k, pad, th = 3, 12, 40, 40, 24, 5, 2, 8.0
    ev = synthetic.events(n, t, h, seed=22)
    kern = np.clip(np.random.default_rng(22).normal(0.8, 0.05, (cout, 2, k, k)), 0, 1).astype(np.float32)
    d = str(tmp_path)
    ev.tofile(os.path.join(d, "events.u8"))
    kern.tofile(os.path.join(d, "k.f32"))
    with open(os.path.join(d, "shape.txt"), "w") as f:
        f.write("%d %d %d %d %d %d %d %.17g\n" % (n, t, h, w, cout, k, pad, th))
    r = subprocess.run([_driver(), "sparse", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "sparse ok" in r.stdout, r.stdout + r.stderr

    def rd(name, dtype, shape):
        return np.fromfile(os.path.join(d, name), dtype=dtype).reshape(shape)

    s = orc.sparse_from_dense(ev)
    out = orc.sparse_inhibit(orc.sparse_conv(s, kern, th, (pad,) * 4))
    dense = orc.sparse_to_dense(out)
    assert dense.any()
    assert np.array_equal(rd("out_dense.out", np.uint8, dense.shape), dense)
    assert np.array_equal(rd("out_gather.out", np.int32, (n, cout, h, w)), orc.sparse_gather(out).astype(np.int32))
    pooled = orc.sparse_to_dense(orc.sparse_pool(out, 2, (1, 1, 1, 1)))
    assert np.array_equal(rd("pooled_dense.out", np.uint8, pooled.shape), pooled)
    ow, oc = orc.sparse_convwta(out, 2, 5)
    want = [[tuple(int(ow[i, j][f]) for f in "ctzyx") for j in range(oc[i])] for i in range(n)]
    assert _winners(os.path.join(d, "winners.out")) == want
    assert np.array_equal(rd("kernel.out", np.float32, kern.shape),
                          orc.sparse_stdp(s, kern, [(0.004, -0.003)], ow, oc, (pad,) * 4))
    fw, fc = orc.fcwta(dense.reshape(n, t, -1), 0, 8)
    fw = fw.reshape(n, 8)
    first = orc.sparse_gather(out).reshape(n, -1)
    got = _winners(os.path.join(d, "fcwta.out"))
    for i in range(n):
        assert [((z * h + y) * w + x, tt) for (_, tt, z, y, x) in got[i]] == \
            [(int(fw[i, j]["x"]), int(first[i, fw[i, j]["x"]])) for j in range(fc[i])]
    times = orc.rank_gather(dense)
    assert np.array_equal(rd("times.out", np.int32, times.shape), times)
    assert np.array_equal(rd("scatter.out", np.uint8, dense.shape), dense)


@pytest.mark.gpu
def test_cpp_front_ends_match_oracle(tmp_path, orc):
    """LoG, ZCA, the Poisson rate code and rate fire through the C++ interface."""
    import ctypes as C
    from spyker_amd import _lib
    n, h, w, rows, cols, t, seed = 6, 28, 28, 800, 48, 15, 4242
    rng = np.random.default_rng(23)
    img = rng.random((n, 1, h, w)).astype(np.float32)
    z = (rng.normal(size=(rows, cols)) @ rng.normal(size=(cols, cols)) * 0.3).astype(np.float32)
    pot = np.cumsum(rng.random((n, t, 4, h, w)).astype(np.float32) * 0.3, axis=1).astype(np.float32)
    d = str(tmp_path)
    img.tofile(os.path.join(d, "img.f32"))
    z.tofile(os.path.join(d, "z.f32"))
    pot.tofile(os.path.join(d, "pot.f32"))
    with open(os.path.join(d, "shape.txt"), "w") as f:
        f.write("%d %d %d %d %d %d %d\n" % (n, h, w, rows, cols, t, seed))
    r = subprocess.run([_driver(), "front", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "front ok" in r.stdout, r.stdout + r.stderr

    def rd(name, dtype, shape):
        return np.fromfile(os.path.join(d, name), dtype=dtype).reshape(shape)

    lk = rd("log_kernel.out", np.float32, (2, 2, 7, 7))
    assert np.array_equal(rd("log.out", np.float32, (n, 4, h, w)), orc.log(img, lk, (3, 3, 3, 3)))
    m64, t64 = orc.zca_fit(z.astype(np.float64), 1e-2)
    m, tr = rd("zca_mean.out", np.float32, (cols,)), rd("zca_trans.out", np.float32, (cols, cols))
    assert (np.abs(m - m64) <= np.spacing(np.abs(m64).astype(np.float32))).all()
    assert (np.abs(tr - t64) <= 4 * 2.0 ** -24 * np.abs(t64) + 1e-9 * np.abs(t64).max()).all()
    white = rd("zca_white.out", np.float32, (rows, cols))
    ref = orc.zca_trans(z, m, tr)
    bound = np.abs(z.astype(np.float64) - m) @ np.abs(tr.astype(np.float64))
    assert (np.abs(white - ref) <= 2.0 ** -24 * np.abs(ref) + 1e-12 * bound).all()
    assert np.array_equal(rd("zca_split.out", np.float32, (rows, 2, cols)), orc.zca_split(white))
    maps = rd("rate_in.out", np.float32, (n, 4 * h * w))
    table = np.zeros((t + 1, 1024), np.uint16)
    assert _lib.load().spk_poisson_table(table.ctypes.data_as(C.c_void_p), C.c_int64(t + 1), C.c_int64(1024)) == 0
    above = min(int((row > row.min()).sum()) for row in maps)
    extra = max(0, min(65536, int((maps.shape[1] - 1) * (t + 1) / max(1, above)) + 2) - (t + 1))
    table = np.concatenate([table, np.repeat(table[-1:], extra, axis=0)])
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    rnd = int(bg.random_raw(1)[0])
    assert np.array_equal(rd("rate.out", np.uint8, (n, t, 4 * h * w)), orc.rate_code(maps, t, True, rnd, table))
    assert np.array_equal(rd("rate_fire.out", np.uint8, pot.shape), orc.rate_fire(pot, 0.25))
