"""CPU-side checks of the C ABI: the library loads, exports what the header
declares, rejects bad arguments before touching a device, and its host-side
kernel generators reproduce the reference's doubles."""

import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spyker_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char \*|int32_t )(spk_\w+)\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    from spyker_amd import _lib
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    from spyker_amd import _lib
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version(lib):
    assert lib.spk_version() == 10000


def _err(lib):
    return lib.spk_last_error().decode()


def test_argument_errors_before_device_work(lib):
    p = C.c_void_p(0)
    assert lib.spk_threshold(p, 9, -1, 1.0, 0.0, p) == -1
    assert "negative" in _err(lib)
    assert lib.spk_threshold(p, 4, 10, 1.0, 0.0, p) == -2  # U8 threshold is not implemented
    assert "U8" in _err(lib)
    assert lib.spk_rank_pool(p, p, 4, 1, 4, 4, 0, 2, 2, 2, 0, 0, 0, 0, p) == -1
    assert "Kernel" in _err(lib)
    assert lib.spk_rank_pool(p, p, 4, 1, 4, 4, 5, 5, 1, 1, 0, 0, 0, 0, p) == -1
    assert "not compatible" in _err(lib)
    assert lib.spk_rank_convwta(p, 9, 1, 1, 1, 1, 1, 0, 0, 0, 0.0, p, p, p) == -1
    assert "Winners count" in _err(lib)
    assert lib.spk_rank_code(p, 9, p, 4, 1, 10, 0, 1, p) == -1
    assert "time" in _err(lib)
    assert lib.spk_conv(p, 4, p, 9, p, 9, 1, 1, 4, 4, 1, 5, 5, 1, 1, 0, 0, 0, 0, p) == -1
    assert "not compatible" in _err(lib)
    from spyker_amd import _lib
    cfg = (_lib.spk_stdp_config * 1)()
    assert lib.spk_rank_convstdp(p, p, 9, p, 1, 1, 1, 4, 4, 1, 3, 3, 2, 2, cfg, 0, p, p, 1, 1, 1, 0, 0, 0, 0, p) == -1
    assert "STDP Config" in _err(lib)
    assert lib.spk_rank_convstdp(p, p, 9, p, 1, 1, 1, 4, 4, 1, 3, 3, 3, 3, cfg, 1, p, p, 1, 1, 1, 0, 0, 0, 0, p) == -1
    assert "not compatible" in _err(lib)


def test_empty_inputs_are_no_ops(lib):
    p = C.c_void_p(0)
    assert lib.spk_threshold(p, 9, 0, 1.0, 0.0, p) == 0
    assert lib.spk_rank_fire(p, 9, p, 4, 0, 0.0, p) == 0
    assert lib.spk_rank_inhibit(p, 9, 0, 3, 4, 16, 0.0, p) == 0
    assert lib.spk_rank_convwta(p, 9, 0, 3, 4, 4, 4, 1, 1, 2, 0.0, p, p, p) == 0


def _host(lib, fn, shape, *args):
    k = np.zeros(shape, np.float64)
    assert getattr(lib, fn)(k.ctypes.data_as(C.c_void_p), *args) == 0
    return k


def test_host_kernel_generators_match_reference(lib, golden):
    import math
    for kk, sd in [(3, 1.0), (3, 2.0), (7, 1.0), (7, 2.0)]:
        assert np.array_equal(_host(lib, "spk_gaussian_kernel", (kk, kk), kk, sd), golden["gauss_%d_%g" % (kk, sd)])
    for th in (0, 45, 90, 135):
        for kk, lam in [(7, 4.0), (11, 8.0)]:
            g = _host(lib, "spk_gabor_kernel", (kk, kk), kk, 0.56 * lam, math.radians(th), 0.5, lam, 0.0)
            assert np.array_equal(g, golden["gabor_%d_%d" % (kk, th)])
    assert lib.spk_random_seed(7) == 0
    assert np.array_equal(_host(lib, "spk_normal_kernel", (64,), 64, 0.8, 0.05), golden["normal_seed7"])


def test_log_kernel_bank(golden):
    """LoG(size, stds) holds the Gaussians of std/sqrt 2 and std*sqrt 2 (opers.cpp:773-788)."""
    import math
    import spyker_amd as spk
    lg = spk.LoG(3, [math.sqrt(2), 2 * math.sqrt(2)], device="cpu", dtype="f64")
    k = lg.kernel.numpy()
    assert k.shape == (2, 2, 7, 7)
    assert np.array_equal(k[0, 0], golden["gauss_7_1"])  # sqrt2 / sqrt2 == 1 exactly
    assert np.array_equal(k[0, 1], golden["gauss_7_2"])  # 2 sqrt2 / sqrt2 == 2 exactly
    np.testing.assert_allclose(k[1, 0], golden["gauss_7_2"], rtol=1e-14, atol=0)
    with pytest.raises(RuntimeError):
        spk.LoG(0, [1.0], device="cpu")
    with pytest.raises(RuntimeError):
        spk.LoG(2, [], device="cpu")


def test_poisson_table_rows_match_reference(lib, golden):
    """spk_poisson_table (Core::poisson_create, cpu/init.cpp:120-149): every row holds the
    reference's intervals; only their order (the reference's std::rand shuffle) differs."""
    t = np.zeros((40, 1024), np.uint16)
    assert lib.spk_poisson_table(t.ctypes.data_as(C.c_void_p), C.c_int64(40), C.c_int64(1024)) == 0
    assert np.array_equal(np.sort(t, axis=1), golden["poisson_rows_sorted_T40"])
    u = np.zeros((41, 1024), np.uint16)
    assert lib.spk_poisson_table(u.ctypes.data_as(C.c_void_p), C.c_int64(41), C.c_int64(1024)) == 0
    assert np.array_equal(u[:40], t)  # a row does not depend on the table's length
    assert (t >= 1).all() and not all((np.diff(r.astype(np.int32)) >= 0).all() for r in t)  # shuffled


def test_python_front_end_shape_rules():
    import spyker_amd as spk
    assert spk.expand2(3) == (3, 3) and spk.expand2([2, 5]) == (2, 5)
    assert spk.expand4(1) == (1, 1, 1, 1) and spk.expand4((1, 2)) == (1, 2, 1, 2)
    assert spk.expand4((1, 2, 3, 4)) == (1, 2, 3, 4)
    with pytest.raises(RuntimeError):
        spk.expand4((1, 2, 3))
    g = spk.GaborFilter(2.0, 90.0, 0.5, 4.0, 0.0)
    assert g.theta == 90.0 * 0.017453292519943295
