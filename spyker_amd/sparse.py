"""Sparse event path on the device (src/python/spyker/sparse.py; src/spyker/opers.cpp:982-1044).

A ``SparseTensor`` of shape [n, t, c, h, w] holds, per sample and step, the
list of spikes that are new at that step (src/spyker/sparse/base.h:493-528),
as device buffers: ``offsets`` (int64, n*t + 1) and ``events`` (int32 packed
z << 16 | y << 8 | x).  Channels, rows and columns are at most 255.  Making a
sparse tensor reads its event count back to the host, so the producing call
synchronizes its stream.

    SparseTensor(tensor, threshold), .dense()   utils.cpp:219-259
    code, conv, gather, pool, inhibit, convwta  sparse.py:37-210
    pad                                         opers.cpp:992-998
    fcwta                                       cpu/wta.cpp:66-126 on .dense()
    Conv.__call__(SparseTensor, threshold)      opers.cpp:866-871
    Conv.stdp(SparseTensor, winners)            opers.cpp:882-889
"""

import ctypes as C

import torch

from . import _lib
from ._lib import call

_GATHER = {"u8": (torch.uint8, _lib.U8), "u16": (torch.uint16, _lib.U16), "i32": (torch.int32, _lib.I32),
           "f32": (torch.float32, _lib.F32)}


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class SparseTensor:
    """Spike lists of a [n, t, c, h, w] tensor on a ROCm device (src/spyker/utils.h:369-420)."""

    def __init__(self, tensor: torch.Tensor = None, threshold: float = 0.0):
        self._s = _lib.spk_sparse()
        self._device = None
        if tensor is None:
            return
        from . import _dev, _type
        _dev(tensor)
        if tensor.dim() != 5:
            raise RuntimeError("Sparse tensors must be 5D.")
        self._device = tensor.device
        n, t, c, h, w = tensor.shape
        call("spk_sparse_from_dense", C.c_void_p(tensor.data_ptr()), _type(tensor), n, t, c, h, w, float(threshold),
             C.byref(self._s), _stream(tensor.device))

    @classmethod
    def _make(cls, device, fn, *args):
        out = cls()
        out._device = device
        call(fn, *args, C.byref(out._s), _stream(device))
        return out

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.offsets:
            try:
                _lib.load().spk_sparse_release(C.byref(s))
            except Exception:
                pass

    def _check(self):
        if not self._s.offsets:
            raise RuntimeError("Sparse can't be used without initialization.")
        return self._s

    def __bool__(self):
        return bool(self._s.offsets)

    @property
    def device(self):
        return self._device

    @property
    def shape(self):
        s = self._s
        return (s.n, s.t, s.c, s.h, s.w) if s.offsets else ()

    @property
    def dims(self):
        return len(self.shape)

    @property
    def numel(self):
        """Number of events (the reference's sparse_elemsize)."""
        return int(self._check().total)

    @property
    def bytes(self):
        s = self._check()
        return int((s.n * s.t + 1) * 8 + s.total * 4)

    def sparsity(self):
        s = self._check()
        return self.numel / float(s.n * s.t * s.c * s.h * s.w)

    def offsets(self) -> torch.Tensor:
        """Copy of the list offsets as an int64 device tensor [n*t + 1]."""
        s = self._check()
        return self._view(s.offsets, s.n * s.t + 1, torch.int64)

    def events(self) -> torch.Tensor:
        """Copy of the packed events as an int32 device tensor [total]."""
        s = self._check()
        return self._view(s.events, s.total, torch.int32)

    def _view(self, ptr, count, dtype):
        out = torch.empty(count, dtype=dtype, device=self._device)
        if count:
            call("spk_copy", C.c_void_p(ptr), C.c_void_p(out.data_ptr()), count * out.element_size(),
                 _stream(self._device))
        return out

    def dense(self) -> torch.Tensor:
        """Cumulative U8 spikes [n, t, c, h, w] (utils.cpp:251-257)."""
        s = self._check()
        out = torch.empty(self.shape, dtype=torch.uint8, device=self._device)
        call("spk_sparse_to_dense", C.byref(s), C.c_void_p(out.data_ptr()), _stream(self._device))
        return out

    def copy(self) -> "SparseTensor":
        s = self._check()
        return SparseTensor._make(self._device, "spk_sparse_pad", C.byref(s), 0, 0, 0, 0)

    def __repr__(self):
        return "SparseTensor(shape=%s, events=%d)" % (self.shape, self._s.total if self._s.offsets else 0)


def _sp(array):
    if not isinstance(array, SparseTensor):
        raise RuntimeError("Expected a SparseTensor.")
    return array._check()


def code(array: torch.Tensor, time: int, sort: bool = True) -> SparseTensor:
    """Rank coding into sparse spike lists (sparse.py:37-64, opers.cpp:1000-1007)."""
    from . import _dev, _type
    _dev(array)
    if array.dim() == 3:
        array = array.unsqueeze(0)
    if array.dim() != 4:
        raise RuntimeError("Input dimensions couldn't be viewed as 4D.")
    if time <= 0:
        raise RuntimeError("Input time can't be less than 1.")
    n, c, h, w = array.shape
    return SparseTensor._make(array.device, "spk_sparse_code", C.c_void_p(array.data_ptr()), _type(array), n, c, h,
                              w, int(time), 1 if sort else 0)


def conv(array: SparseTensor, kernel: torch.Tensor, threshold: float, stride=1, pad=0) -> SparseTensor:
    """Sparse convolution with a dense F32 kernel (sparse.py:67-100, opers.cpp:984-990)."""
    from . import _dev, _type, expand2, expand4
    s = _sp(array)
    _dev(kernel)
    sy, sx = expand2(stride)
    pt, pl, pb, pr = expand4(pad)
    if sy < 1 or sx < 1:
        raise RuntimeError("Stride items are incorrect.")
    if min(pt, pl, pb, pr) < 0:
        raise RuntimeError("Pad items are incorrect.")
    if kernel.dim() != 4 or kernel.shape[1] != s.c:
        raise RuntimeError("Input and kernel shapes are not compatible.")
    cout, _, kh, kw = kernel.shape
    return SparseTensor._make(array.device, "spk_sparse_conv", C.byref(s), C.c_void_p(kernel.data_ptr()),
                              _type(kernel), cout, kh, kw, float(threshold), sy, sx, pt, pl, pb, pr)


def pad(array: SparseTensor, pad) -> SparseTensor:
    """Moves every event by the top/left pad (opers.cpp:992-998)."""
    from . import expand4
    s = _sp(array)
    pt, pl, pb, pr = expand4(pad)
    if min(pt, pl, pb, pr) < 0:
        raise RuntimeError("Pad items are incorrect.")
    return SparseTensor._make(array.device, "spk_sparse_pad", C.byref(s), pt, pl, pb, pr)


def gather(array: SparseTensor, dtype="u8") -> torch.Tensor:
    """First spike step of every element, t where none: [n, c, h, w] (sparse.py:103-125)."""
    s = _sp(array)
    if dtype not in _GATHER:
        raise RuntimeError("Sparse gather: data type %s is not supported." % dtype)
    tt, code_ = _GATHER[dtype]
    out = torch.empty((s.n, s.c, s.h, s.w), dtype=tt, device=array.device)
    call("spk_sparse_gather", C.byref(s), C.c_void_p(out.data_ptr()), code_, _stream(array.device))
    return out


def pool(array: SparseTensor, kernel, stride=None, pad=0) -> SparseTensor:
    """Sparse max pooling, stride = kernel (sparse.py:128-162, opers.cpp:1022-1029)."""
    from . import expand2, expand4
    s = _sp(array)
    kh, kw = expand2(kernel)
    sh, sw = (kh, kw) if stride is None else expand2(stride)
    if (sh, sw) == (0, 0):
        sh, sw = kh, kw
    pt, pl, pb, pr = expand4(pad)
    return SparseTensor._make(array.device, "spk_sparse_pool", C.byref(s), kh, kw, sh, sw, pt, pl, pb, pr)


def inhibit(array: SparseTensor) -> SparseTensor:
    """Each pixel keeps the channel of its first event (sparse.py:165-182)."""
    s = _sp(array)
    return SparseTensor._make(array.device, "spk_sparse_inhibit", C.byref(s))


def convwta(array: SparseTensor, radius, count: int):
    """First-come winner-take-all on sparse lists (sparse.py:185-210, opers.cpp:1038-1043)."""
    from . import Winners, expand2
    s = _sp(array)
    ry, rx = expand2(radius)
    if ry < 0 or rx < 0:
        raise RuntimeError("Radius items are incorrect.")
    if count <= 0:
        raise RuntimeError("Winners count can't be less than 1.")
    data = torch.zeros((s.n, count, 5), dtype=torch.int64, device=array.device)
    cnt = torch.zeros((s.n,), dtype=torch.int64, device=array.device)
    call("spk_sparse_convwta", C.byref(s), ry, rx, int(count), C.c_void_p(data.data_ptr()),
         C.c_void_p(cnt.data_ptr()), _stream(array.device))
    return Winners(data, cnt)


def fcwta(array: SparseTensor, radius: int, count: int):
    """The winners of fcwta(array.dense(), radius, count) (src/spyker/cpu/wta.cpp:66-126) without
    densifying, as conv coordinates: each Winner is (c=0, t=spike step, z, y, x)."""
    from . import Winners
    s = _sp(array)
    if radius < 0:
        raise RuntimeError("Radius can't be less than 0.")
    if count <= 0:
        raise RuntimeError("Winners count can't be less than 1.")
    data = torch.zeros((s.n, count, 5), dtype=torch.int64, device=array.device)
    cnt = torch.zeros((s.n,), dtype=torch.int64, device=array.device)
    call("spk_sparse_fcwta", C.byref(s), int(radius), int(count), C.c_void_p(data.data_ptr()),
         C.c_void_p(cnt.data_ptr()), _stream(array.device))
    return Winners(data, cnt)


def conv_stdp(array: SparseTensor, kernel: torch.Tensor, configs, winners, pad) -> None:
    """STDP of a conv kernel from sparse input spikes (opers.cpp:882-889, sparse/stdp.cpp)."""
    from . import _configs, _type, _winners, expand4
    s = _sp(array)
    if len(winners) == 0:
        raise RuntimeError("Winners can't be empty.")
    cfg, ncfg = _configs(configs)
    wn = _winners(winners, array.device)
    pt, pl, pb, pr = expand4(pad)
    cout, cin, kh, kw = kernel.shape
    if cin != s.c:
        raise RuntimeError("Input and kernel shapes are not compatible.")
    call("spk_sparse_convstdp", C.byref(s), C.c_void_p(kernel.data_ptr()), _type(kernel), cout, kh, kw, cfg, ncfg,
         C.c_void_p(wn.data.data_ptr()), C.c_void_p(wn.counts.data_ptr()), wn.data.shape[1], pt, pl, pb, pr,
         _stream(array.device))
