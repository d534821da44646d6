"""ctypes access to oracle/_ref/libspyker_dropin.so: the reference's own
public C++ API (opers.cpp, impl.cpp, utils.cpp, shape.cpp and its CPU
backend, compiled from the reference's sources by oracle/Makefile) whose CUDA
seam is this repository's HIP backend (spyker_amd/seam/hip_backend.cpp).

Test infrastructure: the functions mirror the pybind lambdas of the
reference's bind.cpp for the layer path, wrapping torch tensors by address as
bind.cpp:184-187 does.
"""

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libspyker_dropin.so")

# torch dtype name -> Spyker Type ordinal (src/spyker/utils.h:96-117)
TYPES = {"int8": 0, "int16": 1, "int32": 2, "int64": 3, "uint8": 4, "float16": 8, "float32": 9, "float64": 10}

WINNER = np.dtype([("c", "<i8"), ("t", "<i8"), ("z", "<i8"), ("y", "<i8"), ("x", "<i8")])


class View(C.Structure):
    _fields_ = [("ptr", C.c_int64), ("device", C.c_int32), ("type", C.c_int32), ("dims", C.c_int32),
                ("pad_", C.c_int32), ("shape", C.c_int64 * 6)]


def available():
    return os.path.exists(PATH)


class DropIn:
    def __init__(self):
        self.lib = C.CDLL(PATH)
        self.lib.dropin_last_error.restype = C.c_char_p

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            raise RuntimeError(self.lib.dropin_last_error().decode())

    @staticmethod
    def view(t):
        """A torch tensor as the reference's wrap of its address (bind.cpp:184-187)."""
        v = View()
        v.ptr = t.data_ptr()
        v.device = t.device.index if t.device.type == "cuda" else -1
        v.type = TYPES[str(t.dtype).replace("torch.", "")]
        v.dims = t.dim()
        for i, s in enumerate(t.shape):
            v.shape[i] = s
        return C.byref(v)

    def device_count(self):
        n = C.c_int64(0)
        self._call("dropin_cuda_device_count", C.byref(n))
        return n.value

    def dog(self, device, size, filters, pad, x, out):
        s1 = (C.c_double * len(filters))(*[f[0] for f in filters])
        s2 = (C.c_double * len(filters))(*[f[1] for f in filters])
        self._call("dropin_dog", C.c_int32(device), C.c_int64(size), s1, s2, C.c_int64(len(filters)),
                   *[C.c_int64(p) for p in pad], self.view(x), self.view(out))

    def gabor(self, device, size, filters, pad, x, out):
        p = (C.c_double * (5 * len(filters)))(*[v for f in filters for v in f])
        self._call("dropin_gabor", C.c_int32(device), C.c_int64(size), p, C.c_int64(len(filters)),
                   *[C.c_int64(q) for q in pad], self.view(x), self.view(out))

    def threshold(self, x, th, value=0.0):
        self._call("dropin_threshold", self.view(x), C.c_double(th), C.c_double(value))

    def code(self, x, out, time, sort=True):
        self._call("dropin_code", self.view(x), self.view(out), C.c_int64(time), C.c_int32(int(sort)))

    def inhibit(self, x, th=0.0):
        self._call("dropin_inhibit", self.view(x), C.c_double(th))

    def convwta(self, x, radius, count, th=0.0):
        n = x.shape[0]
        w = np.zeros((n, count), WINNER)
        c = np.zeros(n, np.int64)
        self._call("dropin_convwta", self.view(x), C.c_int64(radius), C.c_int64(radius), C.c_int64(count),
                   C.c_double(th), w.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
        return w, c

    def fire(self, x, out, th=0.0):
        self._call("dropin_fire", self.view(x), self.view(out), C.c_double(th))

    def pool(self, x, out, kernel, stride, pad=(0, 0, 0, 0)):
        self._call("dropin_pool", self.view(x), self.view(out), C.c_int64(kernel), C.c_int64(kernel),
                   C.c_int64(stride), C.c_int64(stride), *[C.c_int64(p) for p in pad])

    def pad(self, x, out, pad, value=0.0):
        self._call("dropin_pad", self.view(x), self.view(out), *[C.c_int64(p) for p in pad], C.c_double(value))


class Conv:
    """The reference's Conv module on the device (opers.cpp:838-889)."""

    def __init__(self, lib, device, cin, cout, k, stride=1, pad=(0, 0, 0, 0)):
        self.lib = lib
        self.h = C.c_void_p()
        lib._call("dropin_conv_create", C.c_int32(device), C.c_int64(cin), C.c_int64(cout), C.c_int64(k),
                  C.c_int64(k), C.c_int64(stride), C.c_int64(stride), *[C.c_int64(p) for p in pad], C.byref(self.h))

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.lib.dropin_conv_destroy(self.h)

    def set_kernel(self, k):
        self.lib._call("dropin_conv_set_kernel", self.h, self.lib.view(k))

    def add_config(self, positive, negative, stabilize=True, lower=0.0, upper=1.0):
        self.lib._call("dropin_conv_add_config", self.h, C.c_double(positive), C.c_double(negative),
                       C.c_int32(int(stabilize)), C.c_double(lower), C.c_double(upper))

    def forward(self, x, out):
        self.lib._call("dropin_conv_forward", self.h, self.lib.view(x), self.lib.view(out))

    def stdp(self, x, winners, counts, out):
        w = np.ascontiguousarray(winners, WINNER)
        c = np.ascontiguousarray(counts, np.int64)
        self.lib._call("dropin_conv_stdp", self.h, self.lib.view(x), w.ctypes.data_as(C.c_void_p),
                       c.ctypes.data_as(C.c_void_p), C.c_int64(w.shape[0]), C.c_int64(w.shape[1]),
                       self.lib.view(out))
