"""Spyker's dense STDP layer path on MI355X (gfx950).

The functions and modules here mirror the reference's Python package
(src/python/spyker/module.py) for the hot path -- same names, argument
meaning, defaults and shape rules -- over torch tensors on a ROCm device.
Every computation runs in libspyker_hip.so through its C ABI
(include/spyker_hip.h); nothing falls back to the CPU.

    DoG, Gabor, Conv, FC            module.py:174,253,537,665
    LoG, ZCA                        module.py:332,400
    threshold, fire, code           module.py:923,1072,993
    pool, inhibit, convwta, fcwta   module.py:1177,1230,1293,1262
    conv, fc, pad                   module.py:835,801,882
    STDPConfig, DoGFilter, GaborFilter, Winner

Winner lists stay on the device: convwta/fcwta return a ``Winners`` object
holding device tensors, which ``Conv.stdp``/``FC.stdp`` consume directly and
which reads like the reference's list of lists when indexed.
"""

import ctypes as C
import math
from typing import List, Sequence, Union

import numpy as np
import torch

from . import _lib
from ._lib import call

__all__ = [
    "DoGFilter", "GaborFilter", "STDPConfig", "Winner", "Winners", "DoG", "LoG", "ZCA", "Gabor", "Conv", "FC",
    "threshold", "fire", "code", "pool", "inhibit", "convwta", "fcwta", "conv", "fc", "pad", "gather", "scatter", "random_seed",
    "expand2", "expand4",
    "layer_step", "LayerPlan", "sparse", "SparseTensor", "to_sparse",
]

_TYPES = {torch.uint8: _lib.U8, torch.float16: _lib.F16, torch.float32: _lib.F32, torch.float64: _lib.F64,
          torch.int8: _lib.I8, torch.int16: _lib.I16, torch.int32: _lib.I32, torch.int64: _lib.I64,
          torch.uint16: _lib.U16}
_DTYPES = {"u8": torch.uint8, "u16": torch.uint16, "i8": torch.int8, "i16": torch.int16, "i32": torch.int32,
           "i64": torch.int64, "f16": torch.float16, "f32": torch.float32, "f64": torch.float64}


def _dtype(d):
    if isinstance(d, torch.dtype):
        return d
    return _DTYPES[d]


def _type(t):
    try:
        return _TYPES[t.dtype]
    except KeyError:
        raise RuntimeError("Data type %s is not supported." % t.dtype)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _dev(t):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError("spyker_amd operates on ROCm device tensors (device 'cuda').")
    if not t.is_contiguous():
        raise RuntimeError("Input tensor must be contiguous.")
    return t


# -- shape helpers (src/python/spyker/private.py:90-109, src/spyker/opers.cpp:156-227) --

def expand2(v):
    if isinstance(v, int):
        return (v, v)
    v = tuple(int(x) for x in v)
    if len(v) == 1:
        return (v[0], v[0])
    if len(v) != 2:
        raise RuntimeError("Incorrect input shape.")
    return v


def expand4(v):
    """Pad as (top, left, bottom, right): one value for all sides, (vertical, horizontal), or four."""
    if isinstance(v, int):
        return (v, v, v, v)
    v = tuple(int(x) for x in v)
    if len(v) == 1:
        return (v[0],) * 4
    if len(v) == 2:
        return (v[0], v[1], v[0], v[1])
    if len(v) != 4:
        raise RuntimeError("Incorrect input shape.")
    return v


def _least(t, n):
    s = list(t.shape)
    while len(s) < n:
        s.insert(0, 1)
    return t.reshape(s)


def _exact(t, n):
    s = list(t.shape)
    if len(s) == n - 1:
        s.insert(0, 1)
    if len(s) != n:
        raise RuntimeError("Input dimensions couldn't be viewed as %dD." % n)
    return t.reshape(s)


# -- parameter containers ------------------------------------------------------

class DoGFilter:
    """Difference-of-Gaussians parameters (src/spyker/opers.h:162-173)."""

    def __init__(self, std1: float, std2: float):
        self.std1, self.std2 = float(std1), float(std2)


class GaborFilter:
    """Gabor parameters (src/spyker/opers.h:175-211); theta in degrees."""

    def __init__(self, sigma: float, theta: float, gamma: float, lambda_: float, psi: float):
        self.sigma, self.gamma, self.lambda_, self.psi = float(sigma), float(gamma), float(lambda_), float(psi)
        self.theta = float(theta) * 0.017453292519943295


class STDPConfig:
    """STDP parameters (src/spyker/utils.h:505-524)."""

    def __init__(self, positive: float, negative: float, stabilize: bool = True, lower: float = 0.0,
                 upper: float = 1.0):
        self.positive, self.negative = float(positive), float(negative)
        self.stabilize, self.lower, self.upper = bool(stabilize), float(lower), float(upper)

    def _c(self):
        return _lib.spk_stdp_config(self.positive, self.negative, int(self.stabilize), self.lower, self.upper)

    def __repr__(self):
        return "STDPConfig(%g, %g, %s, %g, %g)" % (self.positive, self.negative, self.stabilize, self.lower,
                                                   self.upper)


class Winner:
    """Winner record (src/spyker/utils.h:483-499)."""

    __slots__ = ("c", "t", "z", "y", "x")

    def __init__(self, c=0, t=0, z=0, y=0, x=0):
        self.c, self.t, self.z, self.y, self.x = int(c), int(t), int(z), int(y), int(x)

    def __eq__(self, o):
        return (self.c, self.t, self.z, self.y, self.x) == (o.c, o.t, o.z, o.y, o.x)

    def __repr__(self):
        return "Winner(c=%d, t=%d, z=%d, y=%d, x=%d)" % (self.c, self.t, self.z, self.y, self.x)


class Winners:
    """Per-sample winner lists kept on the device.

    ``data`` is an int64 device tensor [n, count, 5] of (c, t, z, y, x) and
    ``counts`` an int64 device tensor [n].  Indexing or iterating copies them
    to the host once and yields lists of ``Winner``.
    """

    def __init__(self, data: torch.Tensor, counts: torch.Tensor):
        self.data, self.counts = data, counts
        self._host = None

    @staticmethod
    def from_lists(lists, device):
        n = len(lists)
        cap = max([len(w) for w in lists] + [1])
        arr = np.zeros((n, cap, 5), np.int64)
        cnt = np.zeros(n, np.int64)
        for i, ws in enumerate(lists):
            cnt[i] = len(ws)
            for k, w in enumerate(ws):
                arr[i, k] = (w.c, w.t, w.z, w.y, w.x)
        return Winners(torch.from_numpy(arr).to(device), torch.from_numpy(cnt).to(device))

    def tolist(self) -> List[List[Winner]]:
        if self._host is None:
            d, c = self.data.cpu().numpy(), self.counts.cpu().numpy()
            self._host = [[Winner(*d[i, k]) for k in range(int(c[i]))] for i in range(d.shape[0])]
        return self._host

    def set_config(self, index):
        """Sets every winner's STDP configuration index (R-STDP); index: int or per-sample sequence."""
        idx = torch.as_tensor(index, dtype=torch.int64, device=self.data.device)
        if idx.dim() == 0:
            self.data[:, :, 0] = idx
        else:
            self.data[:, :, 0] = idx.reshape(-1, 1)
        self._host = None

    def __len__(self):
        return self.data.shape[0]

    def __getitem__(self, i):
        return self.tolist()[i]

    def __iter__(self):
        return iter(self.tolist())


def _winners(w, device):
    if isinstance(w, Winners):
        return w
    return Winners.from_lists(w, device)


def random_seed(seed: int):
    """Reseeds the host generator used by Conv/FC initialisation (src/spyker/impl.cpp:43)."""
    call("spk_random_seed", C.c_int64(int(seed)))


# -- free functions ------------------------------------------------------------

def threshold(array: torch.Tensor, threshold: float, value: float = 0.0, inplace: bool = True) -> torch.Tensor:
    """x <= threshold -> value (src/spyker/opers.cpp:472-478)."""
    if not inplace:
        array = array.clone()
    _dev(array)
    call("spk_threshold", _ptr(array), _type(array), array.numel(), float(threshold), float(value), _stream(array))
    return array


def _coding(code):
    """Coding type from its name, any case (src/bind/bind.cpp:63-69)."""
    c = str(code).lower()
    if c not in ("rank", "rate"):
        raise RuntimeError("Unknown codeing given.")
    return c


def fire(array: torch.Tensor, threshold: float = 0.0, dtype="u8", code: str = "rank") -> torch.Tensor:
    """Spikes where x > threshold (rank), or spike counts over the steps of
    [N, T, ...] input (rate) (src/spyker/opers.cpp:522-552)."""
    _dev(array)
    out = torch.empty(array.shape, dtype=_dtype(dtype), device=array.device)
    if _coding(code) == "rank":
        call("spk_rank_fire", _ptr(array), _type(array), _ptr(out), _type(out), array.numel(), float(threshold),
             _stream(array))
        return out
    if array.dim() < 3:
        raise RuntimeError("Input dimensions couldn't be viewed as at least 3D.")
    n, t = array.shape[0], array.shape[1]
    call("spk_rate_fire", _ptr(array), _type(array), _ptr(out), _type(out), n, t, math.prod(array.shape[2:]),
         float(threshold), _stream(array))
    return out


def code(array: torch.Tensor, time: int, sort: bool = True, dtype="u8", code: str = "rank") -> torch.Tensor:
    """Rank-order code into cumulative spikes, or rate (Poisson) code into
    spike counts, [N, time, ...] (src/spyker/opers.cpp:488-511)."""
    _dev(array)
    c = _coding(code)
    a = _least(array, 2)
    if time <= 0:
        raise RuntimeError("Input time can't be less than 1.")
    shape = [a.shape[0], int(time)] + list(a.shape[1:])
    out = torch.empty(shape, dtype=_dtype(dtype), device=array.device)
    n = a.shape[0]
    call("spk_rank_code" if c == "rank" else "spk_rate_code", _ptr(a), _type(a), _ptr(out), _type(out), n,
         a.numel() // max(n, 1), int(time), int(bool(sort)), _stream(array))
    return out


def pool(array: torch.Tensor, kernel, stride=None, pad=0) -> torch.Tensor:
    """Max pooling per plane with zero pad (src/spyker/opers.cpp:600-637)."""
    _dev(array)
    if stride is None:
        stride = kernel
    kh, kw = expand2(kernel)
    sh, sw = expand2(stride)
    if (sh, sw) == (0, 0):
        sh, sw = kh, kw
    pt, pl, pb, pr = expand4(pad)
    a = _exact(array, 5)
    h, w = a.shape[3], a.shape[4]
    oh, ow = (h + pt + pb - kh) // sh + 1, (w + pl + pr - kw) // sw + 1
    if kh < 1 or kw < 1 or sh < 1 or sw < 1 or min(pt, pl, pb, pr) < 0 or oh <= 0 or ow <= 0:
        raise RuntimeError("Input and kernel shapes are not compatible.")
    out = torch.empty(list(a.shape[:3]) + [oh, ow], dtype=a.dtype, device=a.device)
    planes = a.shape[0] * a.shape[1] * a.shape[2]
    call("spk_rank_pool", _ptr(a), _ptr(out), _type(a), planes, h, w, kh, kw, sh, sw, pt, pl, pb, pr, _stream(a))
    return out


def inhibit(array: torch.Tensor, threshold: float = 0.0, inplace: bool = True) -> torch.Tensor:
    """Lateral inhibition across channels (src/spyker/opers.cpp:639-646)."""
    if not inplace:
        array = array.clone()
    _dev(array)
    a = _exact(array, 5)
    n, t, c = a.shape[:3]
    call("spk_rank_inhibit", _ptr(a), _type(a), n, t, c, a.shape[3] * a.shape[4], float(threshold), _stream(a))
    return array


def convwta(array: torch.Tensor, radius, count: int, threshold: float = 0.0) -> Winners:
    """Convolutional winner-take-all (src/spyker/opers.cpp:656-662)."""
    _dev(array)
    ry, rx = expand2(radius)
    if ry < 0 or rx < 0:
        raise RuntimeError("Radius items are incorrect.")
    if count <= 0:
        raise RuntimeError("Winners count can't be less than 1.")
    a = _exact(array, 5)
    n, t, c, h, w = a.shape
    data = torch.zeros((n, count, 5), dtype=torch.int64, device=a.device)
    cnt = torch.zeros((n,), dtype=torch.int64, device=a.device)
    call("spk_rank_convwta", _ptr(a), _type(a), n, t, c, h, w, ry, rx, int(count), float(threshold), _ptr(data),
         _ptr(cnt), _stream(a))
    return Winners(data, cnt)


def fcwta(array: torch.Tensor, radius: int, count: int, threshold: float = 0.0) -> Winners:
    """Fully-connected winner-take-all (src/spyker/opers.cpp:648-654)."""
    _dev(array)
    if radius < 0:
        raise RuntimeError("Radius can't be less than 0.")
    if count <= 0:
        raise RuntimeError("Winners count can't be less than 1.")
    a = _exact(array, 3)
    n, t, f = a.shape
    data = torch.zeros((n, count, 5), dtype=torch.int64, device=a.device)
    cnt = torch.zeros((n,), dtype=torch.int64, device=a.device)
    call("spk_rank_fcwta", _ptr(a), _type(a), n, t, f, int(radius), int(count), float(threshold), _ptr(data),
         _ptr(cnt), _stream(a))
    return Winners(data, cnt)


def conv(array: torch.Tensor, kernel: torch.Tensor, stride=1, pad=0) -> torch.Tensor:
    """Time-unrolled cross-correlation, T folded into the batch (src/spyker/opers.cpp:400-423)."""
    _dev(array)
    _dev(kernel)
    sy, sx = expand2(stride)
    pt, pl, pb, pr = expand4(pad)
    a = _exact(array, 5)
    n, t, cin, h, w = a.shape
    cout, kc, kh, kw = kernel.shape
    if kc != cin:
        raise RuntimeError("Input and kernel channels are not compatible.")
    ho, wo = (h + pt + pb - kh) // sy + 1, (w + pl + pr - kw) // sx + 1
    out = torch.empty((n, t, cout, ho, wo), dtype=kernel.dtype, device=a.device)
    call("spk_conv", _ptr(a), _type(a), _ptr(kernel), _type(kernel), _ptr(out), _type(out), n * t, cin, h, w, cout,
         kh, kw, sy, sx, pt, pl, pb, pr, _stream(a))
    return out


def fc(array: torch.Tensor, kernel: torch.Tensor) -> torch.Tensor:
    """out[n,t,o] = sum_f in[n,t,f] k[o,f] (src/spyker/opers.cpp:425-445)."""
    _dev(array)
    _dev(kernel)
    a = _exact(array, 3)
    n, t, fin = a.shape
    fout, kf = kernel.shape
    if kf != fin:
        raise RuntimeError("Input and kernel shapes don't match.")
    out = torch.empty((n, t, fout), dtype=kernel.dtype, device=a.device)
    call("spk_fc", _ptr(a), _type(a), _ptr(kernel), _ptr(out), n * t, fin, fout, _stream(a))
    return out


def gather(array: torch.Tensor, threshold: float = 0.0, dtype="u8", code: str = "rank") -> torch.Tensor:
    """First spike step of every element (rank), or the last step (rate): [N, T, ...] -> [N, ...]
    (src/python/spyker/module.py:1106-1142, src/spyker/opers.cpp:554-575)."""
    _dev(array)
    code = _coding(code)
    a = array if array.dim() >= 3 else array.reshape((1,) * (3 - array.dim()) + tuple(array.shape))
    n, t = a.shape[0], a.shape[1]
    out = torch.empty((n,) + tuple(a.shape[2:]), dtype=_dtype(dtype), device=a.device)
    ln = out[0].numel() if n else 0
    if code == "rank":
        call("spk_rank_gather", _ptr(a), _type(a), _ptr(out), _type(out), n, t, ln, float(threshold), _stream(a))
    else:
        call("spk_rate_gather", _ptr(a), _type(a), _ptr(out), _type(out), n, t, ln, _stream(a))
    return out


def scatter(array: torch.Tensor, time: int, dtype="u8") -> torch.Tensor:
    """Spike trains from first-spike times: [N, ...] -> [N, time, ...], step s spikes where
    s >= the time (src/python/spyker/module.py:1145-1174, src/spyker/opers.cpp:577-598)."""
    _dev(array)
    if time <= 0:
        raise RuntimeError("Input time can't be less than 1.")
    a = array if array.dim() >= 2 else array.reshape(1, -1)
    n = a.shape[0]
    out = torch.empty((n, int(time)) + tuple(a.shape[1:]), dtype=_dtype(dtype), device=a.device)
    ln = a[0].numel() if n else 0
    call("spk_rank_scatter", _ptr(a), _type(a), _ptr(out), _type(out), n, ln, int(time), _stream(a))
    return out


def pad(array: torch.Tensor, pad, value: float = 0.0) -> torch.Tensor:
    """Spatial padding of the last two dimensions (src/spyker/opers.cpp:447-470)."""
    _dev(array)
    pt, pl, pb, pr = expand4(pad)
    a = _exact(array, 5)
    h, w = a.shape[3], a.shape[4]
    out = torch.empty(list(a.shape[:3]) + [h + pt + pb, w + pl + pr], dtype=a.dtype, device=a.device)
    call("spk_pad", _ptr(a), _ptr(out), _type(a), a.shape[0] * a.shape[1] * a.shape[2], h, w, pt, pl, pb, pr,
         float(value), _stream(a))
    return out


def _to_f32(t):
    if t.dtype == torch.float32:
        return t
    out = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    call("spk_cast", _ptr(t), _type(t), _ptr(out), _lib.F32, t.numel(), _stream(t))
    return out


def _host_kernel(fn, shape, *args):
    k = np.zeros(shape, np.float64)
    call(fn, k.ctypes.data_as(C.c_void_p), *args)
    return k


def _device(device):
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


# -- modules -------------------------------------------------------------------

class DoG:
    """Difference-of-Gaussians filter bank (src/spyker/opers.cpp:702-735)."""

    def __init__(self, size: int, filters: Union[DoGFilter, Sequence[DoGFilter]], pad=0, device=None, dtype="f32"):
        if size <= 0:
            raise RuntimeError("Input size can't be less than 1.")
        flist = [filters] if isinstance(filters, DoGFilter) else list(filters)
        if not flist:
            raise RuntimeError("Input filters count can't be less then 1.")
        k = 2 * size + 1
        ker = np.zeros((2, len(flist), k, k), np.float64)
        for i, f in enumerate(flist):
            ker[0, i] = _host_kernel("spk_gaussian_kernel", (k, k), k, f.std1)
            ker[1, i] = _host_kernel("spk_gaussian_kernel", (k, k), k, f.std2)
        self._pad = expand4(pad)
        self.kernel = torch.from_numpy(ker).to(_dtype(dtype)).to(_device(device))

    def __call__(self, array: torch.Tensor) -> torch.Tensor:
        _dev(array)
        a = _to_f32(_exact(array, 4)).contiguous()
        n, c, h, w = a.shape
        count, k = self.kernel.shape[1], self.kernel.shape[2]
        pt, pl, pb, pr = self._pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        if ho <= 0 or wo <= 0:
            raise RuntimeError("Input and kernel shapes are not compatible.")
        out = torch.empty((n, c * count, ho, wo), dtype=self.kernel.dtype, device=a.device)
        call("spk_dog", _ptr(a), _ptr(self.kernel), _ptr(out), n * c, h, w, count, k, pt, pl, pb, pr, _stream(a))
        return out


class LoG:
    """Laplacian-of-Gaussian filter bank (src/spyker/opers.cpp:773-802): each std s gives the
    Gaussian pair (s / sqrt 2, s * sqrt 2); an input channel yields [on, off] x count channels."""

    def __init__(self, size: int, stds: Sequence[float], pad=0, device=None, dtype="f32"):
        if size <= 0:
            raise RuntimeError("Input size can't be less than 1.")
        stds = [float(s) for s in stds]
        if not stds:
            raise RuntimeError("Input stds count can't be less then 1.")
        k = 2 * size + 1
        ker = np.zeros((2, len(stds), k, k), np.float64)
        for i, s in enumerate(stds):
            ker[0, i] = _host_kernel("spk_gaussian_kernel", (k, k), k, s / math.sqrt(2))
            ker[1, i] = _host_kernel("spk_gaussian_kernel", (k, k), k, s * math.sqrt(2))
        self._pad = expand4(pad)
        self.kernel = torch.from_numpy(ker).to(_dtype(dtype)).to(_device(device))

    def __call__(self, array: torch.Tensor) -> torch.Tensor:
        _dev(array)
        a = _to_f32(_exact(array, 4)).contiguous()
        n, c, h, w = a.shape
        count, k = self.kernel.shape[1], self.kernel.shape[2]
        pt, pl, pb, pr = self._pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        if ho <= 0 or wo <= 0:
            raise RuntimeError("Input and kernel shapes are not compatible.")
        if self.kernel.dtype != torch.float32:
            raise RuntimeError("LoG: the kernel must be F32.")
        out = torch.empty((n, c * 2 * count, ho, wo), dtype=torch.float32, device=a.device)
        call("spk_log", _ptr(a), _ptr(self.kernel), _ptr(out), n * c, h, w, count, k, pt, pl, pb, pr, _stream(a))
        return out


def _least2(t):
    if t.dim() == 0:
        raise RuntimeError("Input dimensions couldn't be viewed as at least 2D.")
    return t.reshape(1, -1) if t.dim() == 1 else t


def _zca_input(t):
    _dev(t)
    if t.dtype != torch.float32:
        raise RuntimeError("Input data type must be F32.")
    return t


class ZCA:
    """ZCA whitening (src/spyker/opers.cpp:805-831, src/python/spyker/module.py:400-530).

    ``fit`` learns ``mean`` [features] and ``transform`` [features, features]
    (F32 device tensors, computed in FP64) from the rows of an F32 array;
    calling the module whitens arrays in place (or copies with inplace=False).
    """

    def __init__(self):
        self.mean = None
        self.transform = None

    def fit(self, array: torch.Tensor, epsilon: float, transform: bool = False) -> "ZCA":
        if array.dim() < 2:
            raise RuntimeError("Input dimensions must be at least 2D.")
        a = _zca_input(array)
        rows, cols = a.shape[0], math.prod(a.shape[1:])
        self.mean = torch.empty(cols, dtype=torch.float32, device=a.device)
        self.transform = torch.empty((cols, cols), dtype=torch.float32, device=a.device)
        call("spk_zca_fit", _ptr(a), _ptr(self.mean), _ptr(self.transform), rows, cols, float(epsilon),
             int(bool(transform)), _stream(a))
        return self

    def __call__(self, *array: torch.Tensor, inplace: bool = True):
        if self.mean is None or self.transform is None:
            raise RuntimeError("Layer is not initialized.")
        if not inplace:
            array = [x.clone() for x in array]
        for x in array:
            a = _least2(_zca_input(x))
            rows, cols = a.shape[0], math.prod(a.shape[1:])
            if cols != self.mean.numel() or self.transform.shape != (cols, cols):
                raise RuntimeError("Input and mean shapes are not compatible.")
            if self.mean.device != a.device or self.transform.device != a.device:
                raise RuntimeError("Tensor devices are not compatible.")
            call("spk_zca_trans", _ptr(a), _ptr(self.mean), _ptr(self.transform), rows, cols, _stream(a))
        return tuple(array) if len(array) > 1 else array[0]

    @staticmethod
    def split(array: torch.Tensor) -> torch.Tensor:
        a = _least2(_dev(array))
        shape = list(a.shape)
        shape.insert(1, 2)
        out = torch.empty(shape, dtype=a.dtype, device=a.device)
        call("spk_zca_split", _ptr(a), _ptr(out), _type(a), a.shape[0], math.prod(a.shape[1:]), _stream(a))
        return out

    def save(self, path: str) -> None:
        np.savez(path, mean=self.mean.cpu().numpy(), transform=self.transform.cpu().numpy())

    def load(self, path: str, device=None) -> None:
        data = np.load(path)
        mean, trans = np.asarray(data["mean"], np.float32), np.asarray(data["transform"], np.float32)
        if mean.ndim != 1 or trans.shape != (mean.shape[0], mean.shape[0]):
            raise RuntimeError("Input and mean shapes are not compatible.")
        dev = self.mean.device if self.mean is not None else _device(device)
        self.mean = torch.from_numpy(mean).to(dev)
        self.transform = torch.from_numpy(np.ascontiguousarray(trans)).to(dev)


class Gabor:
    """Gabor filter bank (src/spyker/opers.cpp:739-772)."""

    def __init__(self, size: int, filters: Union[GaborFilter, Sequence[GaborFilter]], pad=0, device=None,
                 dtype="f32"):
        if size <= 0:
            raise RuntimeError("Input size can't be less than 1.")
        flist = [filters] if isinstance(filters, GaborFilter) else list(filters)
        if not flist:
            raise RuntimeError("Input filters count can't be less then 1.")
        k = 2 * size + 1
        ker = np.zeros((len(flist), k, k), np.float64)
        for i, f in enumerate(flist):
            ker[i] = _host_kernel("spk_gabor_kernel", (k, k), k, f.sigma, f.theta, f.gamma, f.lambda_, f.psi)
        self._pad = expand4(pad)
        self.kernel = torch.from_numpy(ker).to(_dtype(dtype)).to(_device(device))

    def __call__(self, array: torch.Tensor) -> torch.Tensor:
        _dev(array)
        a = _to_f32(_exact(array, 4)).contiguous()
        n, c, h, w = a.shape
        count, k = self.kernel.shape[0], self.kernel.shape[1]
        pt, pl, pb, pr = self._pad
        ho, wo = h + pt + pb - k + 1, w + pl + pr - k + 1
        if ho <= 0 or wo <= 0:
            raise RuntimeError("Input and kernel shapes are not compatible.")
        out = torch.empty((n, c * count, ho, wo), dtype=self.kernel.dtype, device=a.device)
        call("spk_gabor", _ptr(a), _ptr(self.kernel), _ptr(out), n * c, h, w, count, k, pt, pl, pb, pr, _stream(a))
        return out


def _configs(lst):
    if not lst:
        raise RuntimeError("At least one STDP Config must be given.")
    arr = (_lib.spk_stdp_config * len(lst))(*[c._c() for c in lst])
    return arr, len(lst)


class Conv:
    """2D convolution layer with rank STDP (src/spyker/opers.cpp:838-889)."""

    def __init__(self, insize: int, outsize: int, kernel, stride=1, pad=0, mean: float = 0.5, std: float = 0.02,
                 device=None, dtype="f32"):
        if insize <= 0:
            raise RuntimeError("Input size can't be less than 1.")
        if outsize <= 0:
            raise RuntimeError("Output size can't be less than 1.")
        kh, kw = expand2(kernel)
        self._stride, self._pad = expand2(stride), expand4(pad)
        ker = _host_kernel("spk_normal_kernel", (outsize, insize, kh, kw), outsize * insize * kh * kw, float(mean),
                           float(std))
        self.kernel = torch.from_numpy(ker).to(_dtype(dtype)).to(_device(device))
        self.stdpconfig: List[STDPConfig] = []

    def __call__(self, array, threshold: float = None):
        if isinstance(array, sparse.SparseTensor):  # opers.cpp:866-871
            return sparse.conv(array, self.kernel, 0.0 if threshold is None else threshold, self._stride, self._pad)
        return conv(array, self.kernel, self._stride, self._pad)

    def stdp(self, array, winners, output: torch.Tensor = None) -> None:
        if isinstance(array, sparse.SparseTensor):  # opers.cpp:882-889
            sparse.conv_stdp(array, self.kernel, self.stdpconfig, winners, self._pad)
            return
        if output is None:
            raise RuntimeError("STDP: dense input needs the output spikes.")
        _dev(array)
        _dev(output)
        if len(winners) == 0:
            raise RuntimeError("Winners can't be empty.")
        cfg, ncfg = _configs(self.stdpconfig)
        a, o = _exact(array, 5), _exact(output, 5)
        if a.dtype != o.dtype:
            raise RuntimeError("Input and output must have the same data type.")
        if a.dtype != torch.uint8:
            raise RuntimeError("STDP: spikes must be u8.")
        wn = _winners(winners, a.device)
        n, t, cin, h, w = a.shape
        cout, _, kh, kw = self.kernel.shape
        sy, sx = self._stride
        pt, pl, pb, pr = self._pad
        ho, wo = o.shape[3], o.shape[4]
        call("spk_rank_convstdp", _ptr(a), _ptr(self.kernel), _type(self.kernel), _ptr(o), n, t, cin, h, w, cout, kh,
             kw, ho, wo, cfg, ncfg, _ptr(wn.data), _ptr(wn.counts), wn.data.shape[1], sy, sx, pt, pl, pb, pr,
             _stream(a))


def layer_step(conv: "Conv", array: torch.Tensor, threshold: float, radius, count: int, pool_kernel=None,
               pool_stride=None, learn: bool = True, potentials: torch.Tensor = None):
    """One training step of a conv layer:

        pot = inhibit(threshold(conv(array), threshold)); winners = convwta(pot, radius, count)
        out = fire(pot); conv.stdp(array, winners, out); pooled = pool(out, pool_kernel, pool_stride)

    On the fused MFMA path (spk_stdp_layer_fused, F32 master weights) in a
    handful of launches that never materialise the FP32 potentials; a layer
    outside that path's domain (F16 or F64 weights, stride other than 1, more
    than 32 steps or 256 output channels, Cin*Kh*Kw above 512, a negative
    threshold, ...) runs the same step on the plain path in the kernel's type
    instead.  Returns (out, winners,
    pooled-or-None).  ``potentials``, if given, is a tensor of the kernel's
    type shaped like ``out`` that receives the raw conv potentials of
    whichever path ran.
    """
    _dev(array)
    a = _exact(array, 5)
    if a.dtype != torch.uint8:
        raise RuntimeError("Fused layer: spikes must be u8.")
    n, t, cin, h, w = a.shape
    cout, _, kh, kw = conv.kernel.shape
    sy, sx = conv._stride
    pt, pl, pb, pr = conv._pad
    ho, wo = (h + pt + pb - kh) // sy + 1, (w + pl + pr - kw) // sx + 1
    ry, rx = expand2(radius)
    pkh = pkw = psh = psw = 0
    if pool_kernel is not None:
        pkh, pkw = expand2(pool_kernel)
        psh, psw = expand2(pool_stride if pool_stride is not None else pool_kernel)
    if potentials is not None:
        _dev(potentials)
        if potentials.dtype != conv.kernel.dtype or tuple(potentials.shape) != (n, t, cout, ho, wo):
            raise RuntimeError("potentials must have the kernel's type and the output's shape")
    if (sy, sx) == (1, 1) and conv.kernel.dtype == torch.float32:
        out = torch.empty((n, t, cout, ho, wo), dtype=torch.uint8, device=a.device)
        pooled = None
        if pool_kernel is not None:
            pooled = torch.empty((n, t, cout, (ho - pkh) // psh + 1, (wo - pkw) // psw + 1), dtype=torch.uint8,
                                 device=a.device)
        # the step writes every winner slot (zeros past a sample's count) and every count
        data = torch.empty((n, count, 5), dtype=torch.int64, device=a.device)
        cnt = torch.empty((n,), dtype=torch.int64, device=a.device)
        cfg, ncfg = _configs(conv.stdpconfig) if learn else (None, 0)
        try:
            call("spk_stdp_layer_fused", _ptr(a), n, t, cin, h, w, _ptr(conv.kernel), cout, kh, kw, pt, pl, pb, pr,
                 float(threshold), ry, rx, int(count), cfg, ncfg, _ptr(out),
                 _ptr(pooled) if pooled is not None else None, pkh, pkw, psh, psw, _ptr(data), _ptr(cnt),
                 _ptr(potentials) if potentials is not None else None, _stream(a))
            return out, Winners(data, cnt), pooled
        except _lib.SpykerError as e:
            if e.code != _lib.ERR_UNSUPPORTED:
                raise
    return _plain_layer_step(conv, a, float(threshold), radius, count, pool_kernel, pool_stride, learn, potentials)


def set_fused_arith(mode: str) -> None:
    """The fused step's whole-sample arithmetic: "int8" (the default: int8 MFMAs on the
    spike bytes, 16-bit fixed-point weights) or "fp16" (FP16 operands)."""
    modes = {"int8": 0, "fp16": 1}
    if mode not in modes:
        raise RuntimeError("Fused arithmetic must be 'int8' or 'fp16'.")
    call("spk_set_fused_arith", modes[mode])


def set_fused_tuning(on: bool) -> None:
    """Whether the fused step times 32 and 64 output channels a wave on the
    first call of each shape and keeps the faster (on by default)."""
    call("spk_set_fused_tuning", 1 if on else 0)


def set_fused_channels(channels: int) -> None:
    """Forces the fused step's output channels a wave (32 or 64; 0 returns to the tuned choice)."""
    call("spk_set_fused_channels", int(channels))


def fused_tuned_channels(conv: "Conv", t: int, h: int, w: int) -> int:
    """The tuned channels a wave for this layer on [*, t, Cin, h, w] input (0 if not tuned yet)."""
    cout, cin, kh, kw = conv.kernel.shape
    pt, pl, pb, pr = conv._pad
    out = C.c_int32(0)
    call("spk_fused_tuned_channels", t, cin, h, w, cout, kh, kw, pt, pl, pb, pr, C.byref(out))
    return out.value


def _plain_layer_step(conv, a, th, radius, count, pool_kernel, pool_stride, learn, potentials):
    """The step as the reference runs it, one plain-path call per stage in the kernel's type."""
    pot = conv(a)
    if potentials is not None:
        potentials.copy_(pot)
    threshold(pot, th)
    inhibit(pot)
    win = convwta(pot, radius, count)
    out = fire(pot)
    del pot
    if learn:
        conv.stdp(a, win, out)
    pooled = pool(out, pool_kernel, pool_stride) if pool_kernel is not None else None
    return out, win, pooled


class LayerPlan:
    """The fused layer step over a fixed batch of instances, each with its own
    size and step count (spk_layer_plan_*): one persistent region kernel per
    step, then convwta, the spike/pool writer and STDP, in the reference's
    batch order (every instance's forward pass, then the winners of instance
    0, 1, ... in rank order).

    ``shapes`` lists (t, h, w) per instance.  Inputs, spikes and pooled spikes
    live in packed U8 buffers (``pack``, ``spikes``, ``pooled``).  The plan
    holds every work buffer, so ``run`` allocates nothing when given its
    output buffers, and a captured run (torch.cuda.graph) replays.
    ``tune=True`` times the region heights that fit on the first run.
    """

    def __init__(self, conv: "Conv", shapes, threshold: float, radius, count: int, pool_kernel=None,
                 pool_stride=None, tune: bool = False):
        sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 3))
        if conv.kernel.dtype != torch.float32:
            raise RuntimeError("Fused layer: the master weights must be f32.")
        if conv._stride != (1, 1):
            raise RuntimeError("Fused layer: stride must be 1.")
        self.conv, self.shapes, self.count = conv, sh, int(count)
        self.device = conv.kernel.device
        cout, cin, kh, kw = conv.kernel.shape
        pt, pl, pb, pr = conv._pad
        ry, rx = expand2(radius)
        pkh = pkw = psh = psw = 0
        if pool_kernel is not None:
            pkh, pkw = expand2(pool_kernel)
            psh, psw = expand2(pool_stride if pool_stride is not None else pool_kernel)
        self._pool = (pkh, pkw, psh, psw)
        n = sh.shape[0]
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            call("spk_layer_plan_create", C.byref(self._h), n, C.c_void_p(sh.ctypes.data), cin, cout, kh, kw, pt, pl,
                 pb, pr, float(threshold), ry, rx, int(count), pkh, pkw, psh, psw, 1 if tune else 0)
        self.in_off, self.out_off, self.pool_off = (np.zeros(n, np.int64) for _ in range(3))
        tot = np.zeros(3, np.int64)
        call("spk_layer_plan_layout", self._h, C.c_void_p(self.in_off.ctypes.data),
             C.c_void_p(self.out_off.ctypes.data), C.c_void_p(self.pool_off.ctypes.data), C.c_void_p(tot.ctypes.data))
        self.in_bytes, self.out_bytes, self.pool_bytes = (int(v) for v in tot)
        self.out_shapes = [(t, cout, h + pt + pb - kh + 1, w + pl + pr - kw + 1) for t, h, w in sh]
        self.pool_shapes = [(t, cout, (ho - pkh) // psh + 1, (wo - pkw) // psw + 1) if pkh else None
                            for t, _, ho, wo in self.out_shapes]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().spk_layer_plan_destroy(h)
            except Exception:
                pass

    def __len__(self):
        return self.shapes.shape[0]

    def pack(self, arrays) -> torch.Tensor:
        """Packs per-instance U8 inputs [t, cin, h, w] into one device buffer."""
        buf = torch.zeros(self.in_bytes, dtype=torch.uint8, device=self.device)
        for i, a in enumerate(arrays):
            flat = a.reshape(-1)
            buf[int(self.in_off[i]):int(self.in_off[i]) + flat.numel()] = flat.to(self.device)
        return buf

    def spikes(self, packed: torch.Tensor, i: int) -> torch.Tensor:
        """Instance i's spikes [t, cout, ho, wo] in a packed spike buffer."""
        shp = self.out_shapes[i]
        return packed[int(self.out_off[i]):int(self.out_off[i]) + int(np.prod(shp))].view(shp)

    def pooled(self, packed: torch.Tensor, i: int) -> torch.Tensor:
        shp = self.pool_shapes[i]
        return packed[int(self.pool_off[i]):int(self.pool_off[i]) + int(np.prod(shp))].view(shp)

    def buffers(self):
        """Fresh (spikes, pooled-or-None, winners data, counts) buffers for run()."""
        out = torch.empty(self.out_bytes, dtype=torch.uint8, device=self.device)
        pooled = torch.empty(self.pool_bytes, dtype=torch.uint8, device=self.device) if self._pool[0] else None
        data = torch.zeros((len(self), self.count, 5), dtype=torch.int64, device=self.device)
        cnt = torch.zeros((len(self),), dtype=torch.int64, device=self.device)
        return out, pooled, data, cnt

    def candidates(self):
        """(rows per region, time chunks) of each candidate tiling, the index runs use, and the
        write mode (1: the region kernel writes the outputs, 0: separate writer kernels)."""
        cap = 16
        rows, chunks = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        count, chosen, writes = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        call("spk_layer_plan_candidates", self._h, C.c_void_p(rows.ctypes.data), C.c_void_p(chunks.ctypes.data), cap,
             C.byref(count), C.byref(chosen), C.byref(writes))
        k = min(count.value, cap)
        return list(zip(rows[:k].tolist(), chunks[:k].tolist())), chosen.value, writes.value

    def select(self, index: int, writes: int = -1) -> None:
        """Runs use candidate ``index`` and write mode ``writes`` (-1 keeps it) from now on."""
        call("spk_layer_plan_select", self._h, int(index), int(writes))

    def run(self, packed: torch.Tensor, learn: bool = True, buffers=None):
        """One step on a packed input; returns (spikes, pooled-or-None, Winners)."""
        _dev(packed)
        if packed.dtype != torch.uint8 or packed.numel() < self.in_bytes:
            raise RuntimeError("Fused layer: packed U8 input of the plan's size expected.")
        out, pooled, data, cnt = buffers if buffers is not None else self.buffers()
        cfg, ncfg = _configs(self.conv.stdpconfig) if learn else (None, 0)
        call("spk_layer_plan_run", self._h, _ptr(packed), _ptr(self.conv.kernel), _type(self.conv.kernel), cfg, ncfg,
             _ptr(out), _ptr(pooled) if pooled is not None else None, _ptr(data), _ptr(cnt), _stream(packed))
        return out, pooled, Winners(data, cnt)


class FC:
    """Fully-connected layer with rank STDP (src/spyker/opers.cpp:891-942)."""

    def __init__(self, insize: int, outsize: int, mean: float = 0.5, std: float = 0.02, device=None, dtype="f32"):
        if insize <= 0:
            raise RuntimeError("Input size can't be less than 1.")
        if outsize <= 0:
            raise RuntimeError("Output size can't be less than 1.")
        ker = _host_kernel("spk_normal_kernel", (outsize, insize), outsize * insize, float(mean), float(std))
        self.kernel = torch.from_numpy(ker).to(_dtype(dtype)).to(_device(device))
        self.stdpconfig: List[STDPConfig] = []

    def __call__(self, array: torch.Tensor) -> torch.Tensor:
        return fc(array, self.kernel)

    def stdp(self, array: torch.Tensor, winners, output: torch.Tensor) -> None:
        _dev(array)
        _dev(output)
        if len(winners) == 0:
            raise RuntimeError("Winners can't be empty.")
        cfg, ncfg = _configs(self.stdpconfig)
        a, o = _exact(array, 3), _exact(output, 3)
        wn = _winners(winners, a.device)
        n, t, fin = a.shape
        call("spk_rank_fcstdp", _ptr(a), _ptr(self.kernel), _type(self.kernel), _ptr(o), n, t, fin,
             self.kernel.shape[0], cfg, ncfg, _ptr(wn.data), _ptr(wn.counts), wn.data.shape[1], _stream(a))


from . import sparse  # noqa: E402
from .sparse import SparseTensor  # noqa: E402


def to_sparse(*arrays: torch.Tensor, threshold: float = 0.0):
    """Dense [n, t, c, h, w] spikes to sparse lists (src/python/spyker/utils.py:506-529)."""
    out = tuple(SparseTensor(a, threshold) for a in arrays)
    return out[0] if len(out) == 1 else out
