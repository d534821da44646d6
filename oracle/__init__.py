"""numpy front end for the CPU oracle and for the reference build.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product
(``spyker_amd``) never imports it.

``Oracle`` wraps ``liboracle.so`` (the C restatement, spyker_oracle.c).
``Reference`` wraps ``_ref/libspyker_ref.so`` (the reference's own CPU
backend compiled from its sources, ref_harness.cpp) when that was built.
Both expose the same methods with the same arguments, so a test can run
one input through each and compare.
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
I64, F32, F64, I32 = C.c_int64, C.c_float, C.c_double, C.c_int32
P = C.c_void_p

WINNER = np.dtype([("c", "<i8"), ("t", "<i8"), ("z", "<i8"), ("y", "<i8"), ("x", "<i8")])
CONFIG = np.dtype(
    [("positive", "<f8"), ("negative", "<f8"), ("stabilize", "<i4"), ("pad", "<i4"), ("lower", "<f8"),
     ("upper", "<f8")],
    align=True)


def build():
    """Compile liboracle.so (and the reference library where its sources exist)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "-j8"])


def _ptr(a):
    return a.ctypes.data_as(P)


def configs_array(configs):
    arr = np.zeros(len(configs), dtype=CONFIG)
    for i, c in enumerate(configs):
        pos, neg = c[0], c[1]
        stab = c[2] if len(c) > 2 else True
        lo = c[3] if len(c) > 3 else 0.0
        hi = c[4] if len(c) > 4 else 1.0
        arr[i] = (pos, neg, int(bool(stab)), 0, lo, hi)
    return arr


class SparseSpikes:
    """Event lists of a [N, T, C, H, W] spike tensor: list (i, j) holds the
    spikes new at step j, as events ev[off[i*T+j]:off[i*T+j+1]] packed
    z << 16 | y << 8 | x (src/spyker/base.h:493-528)."""

    def __init__(self, off, ev, shape):
        self.off, self.ev, self.shape = np.asarray(off, np.int64), np.asarray(ev, np.int32), tuple(shape)

    def lists(self):
        n, t = self.shape[:2]
        return [[[(int(e) >> 16 & 255, int(e) >> 8 & 255, int(e) & 255) for e in self.ev[self.off[i * t + j]:
                                                                                    self.off[i * t + j + 1]]]
                 for j in range(t)] for i in range(n)]

    def __eq__(self, o):
        return self.shape == o.shape and np.array_equal(self.off, o.off) and np.array_equal(self.ev, o.ev)


class _Lib:
    prefix = None
    path = None

    def __init__(self):
        if not os.path.exists(self.path):
            raise FileNotFoundError(self.path + " is not built (run make -C oracle)")
        self.lib = C.CDLL(self.path)

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc, what):
        if rc != 0:
            msg = ""
            if hasattr(self.lib, "ref_last_error"):
                self.lib.ref_last_error.restype = C.c_char_p
                msg = self.lib.ref_last_error().decode()
            raise RuntimeError("%s failed: %s" % (what, msg))

    # rank-order code ------------------------------------------------------
    def rank_code(self, x, time, sort=True):
        """x: [N, L] float32 -> [N, time, L] uint8 (cumulative spikes)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, l = x.shape
        out = np.zeros((n, time, l), np.uint8)
        if self.prefix == "ref_":
            self._check(self._f("rank_code_f32")(_ptr(x), I64(n), I64(l), I64(time), _ptr(out), I32(int(sort))),
                        "rank_code")
        elif sort:
            self._f("rank_code_sort_f32")(_ptr(x), I64(n), I64(l), I64(time), _ptr(out), I32(os.cpu_count() or 1))
        else:
            self._f("rank_code_latency_f32")(_ptr(x), I64(n), I64(l), I64(time), _ptr(out))
        return out

    def threshold(self, x, th, value=0.0, inplace=False):
        """inplace: x (a C-contiguous float32 array) is updated and returned."""
        x = x if inplace else np.array(x, dtype=np.float32, copy=True)
        rc = self._f("threshold_f32")(_ptr(x), I64(x.size), F32(th), F32(value))
        if self.prefix == "ref_":
            self._check(rc, "threshold")
        return x

    def fire(self, x, th=0.0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(x.shape, np.uint8)
        rc = self._f("rank_fire_f32")(_ptr(x), _ptr(out), I64(x.size), F32(th))
        if self.prefix == "ref_":
            self._check(rc, "fire")
        return out

    def inhibit(self, x, th=0.0, inplace=False):
        """x: [N, T, C, H, W] float32 -> inhibited copy (or x itself, updated, when inplace)."""
        x = x if inplace else np.array(x, dtype=np.float32, copy=True)
        n, t, c = x.shape[:3]
        hw = int(np.prod(x.shape[3:]))
        if self.prefix == "ref_":
            self._check(self._f("rank_inhibit_f32")(_ptr(x), I64(n), I64(t), I64(c), I64(hw), F32(th)), "inhibit")
        else:
            self._f("rank_inhibit_f32")(_ptr(x), I64(n), I64(t), I64(c), I64(hw), F32(th), I32(os.cpu_count() or 1))
        return x

    def convwta(self, x, radius, count, th=0.0):
        """x: [N, T, C, H, W] -> list (per sample) of (z, y, x) tuples."""
        ry, rx = (radius, radius) if np.isscalar(radius) else radius
        n, t, c, h, w = x.shape
        win = np.zeros(n * count, WINNER)
        cnt = np.zeros(n, np.int64)
        if x.dtype == np.uint8:
            x = np.ascontiguousarray(x)
            rc = self._f("rank_convwta_u8")(_ptr(x), I64(n), I64(t), I64(c), I64(h), I64(w), I64(ry), I64(rx),
                                             I64(count), _ptr(win), _ptr(cnt))
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            args = [_ptr(x), I64(n), I64(t), I64(c), I64(h), I64(w), I64(ry), I64(rx), I64(count), F32(th),
                    _ptr(win), _ptr(cnt)]
            if self.prefix == "orc_":
                args.append(I32(os.cpu_count() or 1))
            rc = self._f("rank_convwta_f32")(*args)
        if self.prefix == "ref_":
            self._check(rc, "convwta")
        return win.reshape(n, count), cnt

    def fcwta(self, x, radius, count, th=0.0):
        n, t, f = x.shape
        win = np.zeros(n * count, WINNER)
        cnt = np.zeros(n, np.int64)
        if x.dtype == np.uint8:
            x = np.ascontiguousarray(x)
            rc = self._f("rank_fcwta_u8")(_ptr(x), I64(n), I64(t), I64(f), I64(radius), I64(count), _ptr(win),
                                           _ptr(cnt))
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            rc = self._f("rank_fcwta_f32")(_ptr(x), I64(n), I64(t), I64(f), I64(radius), I64(count), F32(th),
                                            _ptr(win), _ptr(cnt))
        if self.prefix == "ref_":
            self._check(rc, "fcwta")
        return win.reshape(n, count), cnt

    def rank_gather(self, x, th=0):
        """x: [N, T, ...] uint8 or float32 -> int32 [N, ...]: the first step whose value is > th
        (th in the input's type), T where none (cpu/time.cpp:40-64)."""
        n, t = x.shape[:2]
        ln = int(np.prod(x.shape[2:]))
        out = np.zeros((n,) + x.shape[2:], np.int32)
        if x.dtype == np.uint8:
            x = np.ascontiguousarray(x)
            rc = self._f("rank_gather_u8")(_ptr(x), I64(n), I64(t), I64(ln), C.c_uint8(int(th)), _ptr(out))
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            rc = self._f("rank_gather_f32")(_ptr(x), I64(n), I64(t), I64(ln), F32(th), _ptr(out))
        if self.prefix == "ref_":
            self._check(rc, "rank_gather")
        return out

    def rank_scatter(self, x, t):
        """x: [N, ...] uint8 or float32 times -> uint8 [N, t, ...]: step s spikes where s >= x
        (cpu/time.cpp:66-84)."""
        n = x.shape[0]
        ln = int(np.prod(x.shape[1:]))
        out = np.zeros((n, t) + x.shape[1:], np.uint8)
        if x.dtype == np.uint8:
            x = np.ascontiguousarray(x)
            rc = self._f("rank_scatter_u8")(_ptr(x), I64(n), I64(ln), I64(t), _ptr(out))
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            rc = self._f("rank_scatter_f32")(_ptr(x), I64(n), I64(ln), I64(t), _ptr(out))
        if self.prefix == "ref_":
            self._check(rc, "rank_scatter")
        return out

    def rate_code(self, x, time, sort, rnd, table):
        """x: [N, L] float32 -> [N, time, L] uint8 spike counts of the Poisson rate code
        (cpu/code.cpp:168-230): sample i draws from std::mt19937(rnd + i), intervals from
        `table` [rows, bins] uint16 at the element's rank-code time row."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        table = np.ascontiguousarray(table, dtype=np.uint16)
        n, l = x.shape
        out = np.zeros((n, time, l), np.uint8)
        if self.prefix == "ref_":
            assert table.shape[1] == 1024
            rc = self._f("rate_code_f32")(_ptr(x), I64(n), I64(l), I64(time), I32(int(sort)), C.c_uint64(rnd),
                                          _ptr(table), I64(table.shape[0]), _ptr(out))
        else:
            rc = self._f("rate_code_f32")(_ptr(x), I64(n), I64(l), I64(time), I32(int(sort)), C.c_uint64(rnd),
                                          _ptr(table), I64(table.shape[0]), I64(table.shape[1]), _ptr(out))
        if rc != 0:
            raise RuntimeError("rate_code: a spike-time row is outside the table")
        return out

    def rate_fire(self, x, threshold):
        """x: [N, T, ...] float32 -> uint8 [N, T, ...] spike counts (cpu/fire.cpp:84-114)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, t = x.shape[:2]
        out = np.zeros(x.shape, np.uint8)
        self._f("rate_fire_f32")(_ptr(x), I64(n), I64(t), I64(int(np.prod(x.shape[2:]))), F32(threshold), _ptr(out))
        return out

    def rate_gather(self, x):
        """x: [N, T, ...] float32 -> float32 [N, ...]: the last step (cpu/time.cpp:86-92)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, t = x.shape[:2]
        out = np.zeros((n,) + x.shape[2:], np.float32)
        rc = self._f("rate_gather_f32")(_ptr(x), I64(n), I64(t), I64(int(np.prod(x.shape[2:]))), _ptr(out))
        if self.prefix == "ref_":
            self._check(rc, "rate_gather")
        return out

    def pad(self, x, pad, value=0):
        """x: [..., H, W] uint8 or float32 -> [..., H+pt+pb, W+pl+pr], the border set to value."""
        pt, pl, pb, pr = pad
        h, w = x.shape[-2:]
        planes = int(np.prod(x.shape[:-2]))
        kind = "u8" if x.dtype == np.uint8 else "f32"
        x = np.ascontiguousarray(x, dtype=np.uint8 if kind == "u8" else np.float32)
        out = np.zeros(x.shape[:-2] + (h + pt + pb, w + pl + pr), x.dtype)
        val = C.c_uint8(int(value)) if kind == "u8" else F32(value)
        rc = self._f("pad_" + kind)(_ptr(x), I64(planes), I64(h), I64(w), _ptr(out), I64(pt), I64(pl), I64(pb),
                                    I64(pr), val)
        if self.prefix == "ref_":
            self._check(rc, "pad")
        return out

    def pool(self, x, kernel, stride=None, pad=(0, 0, 0, 0)):
        """x: [..., H, W] uint8 or float32; pad (top, left, bottom, right)."""
        kh, kw = (kernel, kernel) if np.isscalar(kernel) else kernel
        stride = kernel if stride is None else stride
        sh, sw = (stride, stride) if np.isscalar(stride) else stride
        pt, pl, pb, pr = pad
        h, w = x.shape[-2:]
        planes = int(np.prod(x.shape[:-2]))
        oh, ow = (h + pt + pb - kh) // sh + 1, (w + pl + pr - kw) // sw + 1
        kind = "u8" if x.dtype == np.uint8 else "f32"
        x = np.ascontiguousarray(x, dtype=np.uint8 if kind == "u8" else np.float32)
        out = np.zeros(x.shape[:-2] + (oh, ow), x.dtype)
        rc = self._f("rank_pool_" + kind)(_ptr(x), I64(planes), I64(h), I64(w), _ptr(out), I64(kh), I64(kw),
                                          I64(sh), I64(sw), I64(pt), I64(pl), I64(pb), I64(pr))
        if self.prefix == "ref_":
            self._check(rc, "pool")
        return out

    def convstdp(self, spk_in, kernel, spk_out, configs, winners, counts, stride=1, pad=(0, 0, 0, 0)):
        """Returns the updated kernel (a copy)."""
        spk_in = np.ascontiguousarray(spk_in, np.uint8)
        spk_out = np.ascontiguousarray(spk_out, np.uint8)
        k = np.array(kernel, dtype=np.float32, copy=True)
        cfg = configs_array(configs)
        win = np.ascontiguousarray(winners, WINNER)
        cnt = np.ascontiguousarray(counts, np.int64)
        n, t, cin, h, w = spk_in.shape
        cout, _, kh, kw = k.shape
        ho, wo = spk_out.shape[-2:]
        sy, sx = (stride, stride) if np.isscalar(stride) else stride
        rc = self._f("rank_convstdp_f32")(
            _ptr(spk_in), I64(n), I64(t), I64(cin), I64(h), I64(w), _ptr(k), I64(cout), I64(kh), I64(kw),
            _ptr(spk_out), I64(ho), I64(wo), _ptr(cfg), I64(len(cfg)), _ptr(win), _ptr(cnt), I64(win.shape[1]),
            I64(sy), I64(sx), *[I64(p) for p in pad])
        self._check(rc, "convstdp")
        return k

    def fcstdp(self, spk_in, kernel, spk_out, configs, winners, counts):
        spk_in = np.ascontiguousarray(spk_in, np.uint8)
        spk_out = np.ascontiguousarray(spk_out, np.uint8)
        k = np.array(kernel, dtype=np.float32, copy=True)
        cfg = configs_array(configs)
        win = np.ascontiguousarray(winners, WINNER)
        cnt = np.ascontiguousarray(counts, np.int64)
        n, t, fin = spk_in.shape
        fout = k.shape[0]
        rc = self._f("rank_fcstdp_f32")(_ptr(spk_in), I64(n), I64(t), I64(fin), _ptr(k), I64(fout), _ptr(spk_out),
                                        _ptr(cfg), I64(len(cfg)), _ptr(win), _ptr(cnt), I64(win.shape[1]))
        self._check(rc, "fcstdp")
        return k

    # -- sparse event path -------------------------------------------------

    def _sparse_out(self, name, n, t, cap, *args):
        off = np.zeros(n * t + 1, np.int64)
        ev = np.zeros(max(cap, 1), np.int32)
        f = self._f(name)
        f.restype = C.c_int64
        total = f(*args, _ptr(off), _ptr(ev), I64(cap))
        if total < 0:
            raise RuntimeError("%s: %s" % (name, "capacity exceeded" if total == -1 else "reference error"))
        return off, ev[:total].copy()

    def sparse_code(self, x, time, sort=True):
        """x: [N, C, H, W] float32 -> SparseSpikes of shape [N, time, C, H, W]."""
        x = np.ascontiguousarray(x, np.float32)
        n, c, h, w = x.shape
        off, ev = self._sparse_out("sparse_code_f32", n, time, n * c * h * w, _ptr(x), I64(n), I64(c), I64(h),
                                   I64(w), I64(time), I32(int(bool(sort))))
        return SparseSpikes(off, ev, (n, time, c, h, w))

    def sparse_from_dense(self, x, threshold=0):
        """x: [N, T, C, H, W] uint8 -> SparseSpikes (event at the first step > threshold)."""
        x = np.ascontiguousarray(x, np.uint8)
        n, t, c, h, w = x.shape
        off, ev = self._sparse_out("sparse_from_dense_u8", n, t, n * c * h * w, _ptr(x), I64(n), I64(t), I64(c),
                                   I64(h), I64(w), C.c_uint8(int(threshold)))
        return SparseSpikes(off, ev, x.shape)

    def sparse_to_dense(self, sp):
        n, t, c, h, w = sp.shape
        out = np.zeros(sp.shape, np.uint8)
        self._f("sparse_to_dense_u8")(_ptr(sp.off), _ptr(sp.ev), I64(n), I64(t), I64(c), I64(h), I64(w), _ptr(out))
        return out

    def sparse_gather(self, sp):
        n, t, c, h, w = sp.shape
        out = np.zeros((n, c, h, w), np.uint8)
        self._f("sparse_gather_u8")(_ptr(sp.off), _ptr(sp.ev), I64(n), I64(t), I64(c), I64(h), I64(w), _ptr(out))
        return out

    def sparse_pad(self, sp, pad):
        pt, pl, pb, pr = pad
        n, t, c, h, w = sp.shape
        out = np.zeros(max(len(sp.ev), 1), np.int32)
        self._f("sparse_pad")(_ptr(sp.off), _ptr(sp.ev), I64(n), I64(t), I64(pt), I64(pl), _ptr(out))
        return SparseSpikes(sp.off.copy(), out[:len(sp.ev)].copy(), (n, t, c, h + pt + pb, w + pl + pr))

    def sparse_pool(self, sp, kernel, pad=(0, 0, 0, 0)):
        kh, kw = (kernel, kernel) if np.isscalar(kernel) else kernel
        pt, pl, pb, pr = pad
        n, t, c, h, w = sp.shape
        ho, wo = (h + pt + pb - kh) // kh + 1, (w + pl + pr - kw) // kw + 1
        off, ev = self._sparse_out("sparse_pool", n, t, max(len(sp.ev), 1), _ptr(sp.off), _ptr(sp.ev), I64(n),
                                   I64(t), I64(c), I64(ho), I64(wo), I64(kh), I64(kw), I64(pt), I64(pl))
        return SparseSpikes(off, ev, (n, t, c, ho, wo))

    def sparse_inhibit(self, sp):
        n, t, c, h, w = sp.shape
        off, ev = self._sparse_out("sparse_inhibit", n, t, max(len(sp.ev), 1), _ptr(sp.off), _ptr(sp.ev), I64(n),
                                   I64(t), I64(h), I64(w))
        return SparseSpikes(off, ev, sp.shape)

    def sparse_convwta(self, sp, radius, count):
        ry, rx = (radius, radius) if np.isscalar(radius) else radius
        n, t, c, h, w = sp.shape
        win = np.zeros(n * count, WINNER)
        cnt = np.zeros(n, np.int64)
        self._f("sparse_convwta")(_ptr(sp.off), _ptr(sp.ev), I64(n), I64(t), I64(h), I64(w), I64(ry), I64(rx),
                                  I64(count), _ptr(win), _ptr(cnt))
        return win.reshape(n, count), cnt

    def sparse_conv(self, sp, kernel, threshold, pad=(0, 0, 0, 0)):
        kernel = np.ascontiguousarray(kernel, np.float32)
        pt, pl, pb, pr = pad
        n, t, cin, h, w = sp.shape
        cout, _, kh, kw = kernel.shape
        ho, wo = h + pt + pb - kh + 1, w + pl + pr - kw + 1
        off, ev = self._sparse_out("sparse_conv", n, t, n * cout * ho * wo, _ptr(sp.off), _ptr(sp.ev), I64(n),
                                   I64(t), I64(cin), I64(h), I64(w), _ptr(kernel), I64(cout), I64(kh), I64(kw),
                                   F32(threshold), I64(pt), I64(pl), I64(pb), I64(pr))
        return SparseSpikes(off, ev, (n, t, cout, ho, wo))

    def sparse_stdp(self, sp, kernel, configs, winners, counts, pad=(0, 0, 0, 0)):
        """Returns the updated kernel (a copy); pad (top, left, ...) as given to the conv."""
        k = np.array(kernel, dtype=np.float32, copy=True)
        cfg = configs_array(configs)
        win = np.ascontiguousarray(winners, WINNER)
        cnt = np.ascontiguousarray(counts, np.int64)
        n, t, cin, h, w = sp.shape
        cout, _, kh, kw = k.shape
        rc = self._f("sparse_stdp")(_ptr(sp.off), _ptr(sp.ev), I64(n), I64(t), I64(cin), _ptr(k), I64(cout), I64(kh),
                                    I64(kw), I64(pad[0]), I64(pad[1]), _ptr(cfg), I64(len(cfg)), _ptr(win),
                                    _ptr(cnt), I64(win.shape[1]))
        self._check(rc, "sparse_stdp")
        return k

    def gaussian_kernel(self, k, std):
        out = np.zeros((k, k), np.float64)
        rc = self._f("gaussian_kernel")(_ptr(out), I64(k), F64(std))
        if self.prefix == "ref_":
            self._check(rc, "gaussian_kernel")
        return out

    def gabor_kernel(self, k, sigma, theta, gamma, lam, psi):
        out = np.zeros((k, k), np.float64)
        rc = self._f("gabor_kernel")(_ptr(out), I64(k), F64(sigma), F64(theta), F64(gamma), F64(lam), F64(psi))
        if self.prefix == "ref_":
            self._check(rc, "gabor_kernel")
        return out


class Oracle(_Lib):
    prefix = "orc_"
    path = os.path.join(HERE, "liboracle.so")

    def __init__(self):
        super().__init__()
        self.lib.orc_stdp_update.restype = C.c_float

    def mt_stream(self, seed, n):
        """The first n outputs of std::mt19937(seed)."""
        out = np.zeros(n, np.uint32)
        self.lib.orc_mt_stream(C.c_uint32(seed), I64(n), _ptr(out))
        return out

    def stdp_update(self, w, potentiate, config):
        cfg = configs_array([config])
        return self.lib.orc_stdp_update(F32(w), I32(int(potentiate)), _ptr(cfg))

    def conv(self, x, kernel, stride=1, pad=(0, 0, 0, 0), f64=False, threads=None):
        """x: [N, Cin, H, W] uint8/float32; kernel [Cout, Cin, Kh, Kw] float32."""
        is_u8 = x.dtype == np.uint8
        x = np.ascontiguousarray(x, np.uint8 if is_u8 else np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        n, cin, h, w = x.shape
        cout, _, kh, kw = kernel.shape
        sy, sx = (stride, stride) if np.isscalar(stride) else stride
        pt, pl, pb, pr = pad
        ho, wo = (h + pt + pb - kh) // sy + 1, (w + pl + pr - kw) // sx + 1
        out = np.zeros((n, cout, ho, wo), np.float64 if f64 else np.float32)
        fn = self.lib.orc_conv_f64 if f64 else self.lib.orc_conv_f32
        fn(_ptr(x), I32(int(is_u8)), I64(n), I64(cin), I64(h), I64(w), _ptr(kernel), I64(cout), I64(kh), I64(kw),
           I64(sy), I64(sx), I64(pt), I64(pl), I64(pb), I64(pr), _ptr(out), I32(threads or os.cpu_count() or 1))
        return out

    def conv_f64k(self, x, kernel, stride=1, pad=(0, 0, 0, 0), threads=None):
        """F64 kernel: x [N, Cin, H, W] uint8/float32/float64 converted to F64, F64 fma chain."""
        kind = {np.dtype(np.uint8): 0, np.dtype(np.float32): 1, np.dtype(np.float64): 2}[x.dtype]
        x = np.ascontiguousarray(x)
        kernel = np.ascontiguousarray(kernel, np.float64)
        n, cin, h, w = x.shape
        cout, _, kh, kw = kernel.shape
        sy, sx = (stride, stride) if np.isscalar(stride) else stride
        pt, pl, pb, pr = pad
        ho, wo = (h + pt + pb - kh) // sy + 1, (w + pl + pr - kw) // sx + 1
        out = np.zeros((n, cout, ho, wo), np.float64)
        self.lib.orc_conv_f64k(_ptr(x), I32(kind), I64(n), I64(cin), I64(h), I64(w), _ptr(kernel), I64(cout),
                               I64(kh), I64(kw), I64(sy), I64(sx), I64(pt), I64(pl), I64(pb), I64(pr), _ptr(out),
                               I32(threads or os.cpu_count() or 1))
        return out

    def fc(self, x, kernel):
        is_u8 = x.dtype == np.uint8
        x = np.ascontiguousarray(x, np.uint8 if is_u8 else np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        fin = x.shape[-1]
        rows = x.size // fin
        fout = kernel.shape[0]
        out = np.zeros(x.shape[:-1] + (fout,), np.float32)
        self.lib.orc_fc_f32(_ptr(x), I32(int(is_u8)), I64(rows), I64(fin), _ptr(kernel), I64(fout), _ptr(out))
        return out

    def dog(self, x, kernel, pad):
        """x: [N, C, H, W] float32, kernel [2, count, k, k] float32 -> [N, C*count, Ho, Wo]."""
        x = np.ascontiguousarray(x, np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        n, c, h, w = x.shape
        count, k = kernel.shape[1], kernel.shape[2]
        pt, pl, pb, pr = pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        out = np.zeros((n, c * count, ho, wo), np.float32)
        self.lib.orc_dog_f32(_ptr(x), I64(n * c), I64(h), I64(w), _ptr(kernel), I64(count), I64(k), I64(pt),
                             I64(pl), I64(pb), I64(pr), _ptr(out), I32(os.cpu_count() or 1))
        return out

    def log(self, x, kernel, pad):
        """x: [N, C, H, W] float32, kernel [2, count, k, k] float32 -> [N, C*2*count, Ho, Wo]."""
        x = np.ascontiguousarray(x, np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        n, c, h, w = x.shape
        count, k = kernel.shape[1], kernel.shape[2]
        pt, pl, pb, pr = pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        out = np.zeros((n, c * 2 * count, ho, wo), np.float32)
        self.lib.orc_log_f32(_ptr(x), I64(n * c), I64(h), I64(w), _ptr(kernel), I64(count), I64(k), I64(pt),
                             I64(pl), I64(pb), I64(pr), _ptr(out), I32(os.cpu_count() or 1))
        return out

    # -- ZCA (cpu/zca.cpp:66-156), numpy FP64 restatement ------------------
    # The reference computes in its F32 through BLAS/LAPACK and is not built
    # here (no BLAS/LAPACK in its build recipe); these are the same steps in
    # FP64, pinned by the mathematics in tests/test_oracle.py.

    @staticmethod
    def zca_fit(x, epsilon):
        """x: [rows, cols] -> (mean FP64 [cols], trans FP64 [cols, cols]), zca.cpp:66-97:
        mean of the rows; sigma = mid^T mid / (rows - 1); trans = U diag(1 / sqrt(max(0, l) + eps)) U^T,
        eps taken as the F32 value."""
        x = np.asarray(x, np.float64).reshape(x.shape[0], -1)
        mean = x.mean(axis=0)
        mid = x - mean
        sigma = mid.T @ mid / (x.shape[0] - 1)
        lam, u = np.linalg.eigh(sigma)
        s = 1.0 / np.sqrt(np.maximum(0.0, lam) + float(np.float32(epsilon)))
        return mean, (u * s) @ u.T

    @staticmethod
    def zca_trans(x, mean, trans):
        """(x - mean) . T in FP64, T the symmetric matrix of trans's row-major upper
        triangle (zca.cpp:107-121, rusym)."""
        x = np.asarray(x, np.float64).reshape(x.shape[0], -1)
        t = np.triu(np.asarray(trans, np.float64))
        t = t + np.triu(t, 1).T
        return (x - np.asarray(mean, np.float64)) @ t

    @staticmethod
    def zca_split(x):
        """[rows, ...] -> [rows, 2, ...] of max(x, 0) and max(-x, 0) in the element type
        (zca.cpp:141-156; -x wraps for the integer types as C++'s conversion does)."""
        x = np.asarray(x)
        with np.errstate(over="ignore"):
            neg = (-x).astype(x.dtype)
        zero = x.dtype.type(0)
        return np.stack([np.where(x < zero, zero, x), np.where(neg < zero, zero, neg)], axis=1)

    def gabor(self, x, kernel, pad):
        x = np.ascontiguousarray(x, np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        n, c, h, w = x.shape
        count, k = kernel.shape[0], kernel.shape[1]
        pt, pl, pb, pr = pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        out = np.zeros((n, c * count, ho, wo), np.float32)
        self.lib.orc_gabor_f32(_ptr(x), I64(n * c), I64(h), I64(w), _ptr(kernel), I64(count), I64(k), I64(pt),
                               I64(pl), I64(pb), I64(pr), _ptr(out), I32(os.cpu_count() or 1))
        return out


class Reference(_Lib):
    prefix = "ref_"
    path = os.path.join(HERE, "_ref", "libspyker_ref.so")

    def normal_kernel(self, n, mean, std, seed):
        out = np.zeros(n, np.float64)
        self._check(self.lib.ref_normal_kernel(_ptr(out), I64(n), F64(mean), F64(std), I64(seed)), "normal_kernel")
        return out

    def poisson_create(self, rows, bins=1024):
        """Core::poisson_create (cpu/init.cpp:120-149): [rows, bins] uint16, rows in the
        reference's (std::rand-shuffled) order."""
        out = np.zeros((rows, bins), np.uint16)
        self._check(self.lib.ref_poisson_create(I64(rows), I64(bins), _ptr(out)), "poisson_create")
        return out



# Spyker Type ordinals (src/spyker/utils.h:96-117) of the real numpy types
REF_TYPE = {np.dtype(np.float16): 8, np.dtype(np.float32): 9, np.dtype(np.float64): 10}


class ReferenceTyped(Reference):
    """The reference's own CPU stages for F16, F32 or F64 values (ref_*_t)."""

    def threshold(self, x, th, value=0.0):
        x = np.array(x, copy=True)
        self._check(self.lib.ref_threshold_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(x.size), F64(th), F64(value)),
                    "threshold")
        return x

    def fire(self, x, th=0.0):
        x = np.ascontiguousarray(x)
        out = np.zeros(x.shape, np.uint8)
        self._check(self.lib.ref_rank_fire_t(_ptr(x), I32(REF_TYPE[x.dtype]), _ptr(out), I64(x.size), F64(th)),
                    "fire")
        return out

    def inhibit(self, x, th=0.0):
        x = np.array(x, copy=True)
        n, t, c = x.shape[:3]
        hw = int(np.prod(x.shape[3:]))
        self._check(self.lib.ref_rank_inhibit_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(n), I64(t), I64(c), I64(hw),
                                                F64(th)), "inhibit")
        return x

    def convwta(self, x, radius, count, th=0.0):
        ry, rx = (radius, radius) if np.isscalar(radius) else radius
        x = np.ascontiguousarray(x)
        n, t, c, h, w = x.shape
        win = np.zeros(n * count, WINNER)
        cnt = np.zeros(n, np.int64)
        self._check(self.lib.ref_rank_convwta_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(n), I64(t), I64(c), I64(h),
                                                I64(w), I64(ry), I64(rx), I64(count), F64(th), _ptr(win), _ptr(cnt)),
                    "convwta")
        return win.reshape(n, count), cnt

    def fcwta(self, x, radius, count, th=0.0):
        x = np.ascontiguousarray(x)
        n, t, f = x.shape
        win = np.zeros(n * count, WINNER)
        cnt = np.zeros(n, np.int64)
        self._check(self.lib.ref_rank_fcwta_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(n), I64(t), I64(f), I64(radius),
                                              I64(count), F64(th), _ptr(win), _ptr(cnt)), "fcwta")
        return win.reshape(n, count), cnt

    def pool(self, x, kernel, stride=None, pad=(0, 0, 0, 0)):
        x = np.ascontiguousarray(x)
        kh, kw = (kernel, kernel) if np.isscalar(kernel) else kernel
        stride = kernel if stride is None else stride
        sh, sw = (stride, stride) if np.isscalar(stride) else stride
        pt, pl, pb, pr = pad
        h, w = x.shape[-2:]
        oh, ow = (h + pt + pb - kh) // sh + 1, (w + pl + pr - kw) // sw + 1
        out = np.zeros(x.shape[:-2] + (oh, ow), x.dtype)
        planes = int(np.prod(x.shape[:-2]))
        self._check(self.lib.ref_rank_pool_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(planes), I64(h), I64(w), _ptr(out),
                                             I64(kh), I64(kw), I64(sh), I64(sw), I64(pt), I64(pl), I64(pb), I64(pr)),
                    "pool")
        return out

    def pad(self, x, pad, value=0.0):
        x = np.ascontiguousarray(x)
        pt, pl, pb, pr = pad
        h, w = x.shape[-2:]
        out = np.zeros(x.shape[:-2] + (h + pt + pb, w + pl + pr), x.dtype)
        planes = int(np.prod(x.shape[:-2]))
        self._check(self.lib.ref_pad_t(_ptr(x), I32(REF_TYPE[x.dtype]), I64(planes), I64(h), I64(w), _ptr(out),
                                       I64(pt), I64(pl), I64(pb), I64(pr), F64(value)), "pad")
        return out

    def convstdp(self, spk_in, kernel, spk_out, configs, winners, counts, stride=1, pad=(0, 0, 0, 0)):
        spk_in = np.ascontiguousarray(spk_in, np.uint8)
        spk_out = np.ascontiguousarray(spk_out, np.uint8)
        k = np.array(kernel, copy=True)
        cfg = configs_array(configs)
        win = np.ascontiguousarray(winners, WINNER)
        cnt = np.ascontiguousarray(counts, np.int64)
        n, t, cin, h, w = spk_in.shape
        cout, _, kh, kw = k.shape
        ho, wo = spk_out.shape[-2:]
        sy, sx = (stride, stride) if np.isscalar(stride) else stride
        rc = self.lib.ref_rank_convstdp_t(
            _ptr(spk_in), I64(n), I64(t), I64(cin), I64(h), I64(w), _ptr(k), I32(REF_TYPE[k.dtype]), I64(cout),
            I64(kh), I64(kw), _ptr(spk_out), I64(ho), I64(wo), _ptr(cfg), I64(len(cfg)), _ptr(win), _ptr(cnt),
            I64(win.shape[1]), I64(sy), I64(sx), *[I64(p) for p in pad])
        self._check(rc, "convstdp")
        return k


def reference_available():
    return os.path.exists(Reference.path)
