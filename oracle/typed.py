"""numpy restatement of the plain path's typed stages (F16, F32, F64 values).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product
(``spyker_amd``) never imports it.

The reference computes these stages in the tensor's own type T, F16 through
half_float::half (3rd/half, round to nearest, every operation correctly
rounded); numpy's float16 arithmetic rounds each operation the same way (a
float16 sum, difference, product or quotient computed in float32 and rounded
once is the correctly rounded float16 result).  Small inputs only: the loops
are per sample and per winner.

Pinned two ways (tests/test_oracle_typed.py): at T = float32 every function
equals the C oracle (itself pinned to the reference's golden vectors), and
for F16 and F64 every function equals the reference's own CPU code compiled
by oracle/Makefile (oracle/_ref, ref_*_t entry points) where that is built.
"""

from fractions import Fraction

import numpy as np

WINNER = np.dtype([("c", "<i8"), ("t", "<i8"), ("z", "<i8"), ("y", "<i8"), ("x", "<i8")])

# Limits<T>::min() (src/spyker/base.h:761-774): the lowest F32/F64, but
# numeric_limits<half>::min() -- the smallest positive normal, 2^-14 -- for F16
_LOWEST = {np.float16: np.float16(2.0 ** -14), np.float32: np.float32(-np.finfo(np.float32).max),
           np.float64: np.float64(-np.finfo(np.float64).max)}
_EPSILON = np.float32(1e-5)  # base.h:776, a float


def lowest(dt):
    return _LOWEST[np.dtype(dt).type]


def threshold(x, th, value=0.0):
    """cpu/threshold: x <= th -> value, in x's type."""
    t = x.dtype.type
    x = x.copy()
    x[x <= t(th)] = t(value)
    return x


def fire(x, th=0.0):
    """cpu/fire.cpp rank: x > th (th in x's type) -> 1."""
    return (x > x.dtype.type(th)).astype(np.uint8)


def inhibit(x, th=0.0):
    """cpu/inhibit.cpp:41-121 on [N, T, C, H, W] in x's type: the channel
    maximum per step (first maximum), the count of steps whose maximum
    exceeds th, the kept channel at step T - max(cnt, 1), every value times
    (channel == kept)."""
    dt = x.dtype.type
    n, t, c = x.shape[:3]
    xs = x.reshape(n, t, c, -1).copy()
    m = xs.max(axis=2)                              # [n, t, hw]
    cnt = (m > dt(th)).sum(axis=1)                  # [n, hw]
    ts = t - np.maximum(cnt, 1)                     # [n, hw]
    hw = xs.shape[3]
    rows = xs[np.arange(n)[:, None], ts, :, np.arange(hw)[None, :]]  # [n, hw, c]
    keep = rows.argmax(axis=2)                      # first maximum
    sel = (np.arange(c)[None, :, None] == keep[:, None, :]).astype(dt)  # [n, c, hw]
    xs = xs * sel[:, None, :, :]
    return xs.reshape(x.shape)


def pool(x, kernel, stride=None, pad=(0, 0, 0, 0)):
    """cpu/pool.cpp:53-86 on [..., H, W]: zero pad, then each window's maximum
    in row-major scan order (std::max keeps the earlier of equals) seeded with
    Limits<T>::min() -- for F16 the smallest positive normal."""
    dt = x.dtype.type
    kh, kw = (kernel, kernel) if np.isscalar(kernel) else kernel
    stride = kernel if stride is None else stride
    sh, sw = (stride, stride) if np.isscalar(stride) else stride
    pt, pl, pb, pr = pad
    xp = np.pad(x, [(0, 0)] * (x.ndim - 2) + [(pt, pb), (pl, pr)])
    h, w = xp.shape[-2:]
    oh, ow = (h - kh) // sh + 1, (w - kw) // sw + 1
    out = np.full(x.shape[:-2] + (oh, ow), lowest(dt), dt)
    for a in range(kh):
        for b in range(kw):
            v = xp[..., a:a + (oh - 1) * sh + 1:sh, b:b + (ow - 1) * sw + 1:sw]
            out = np.where(out < v, v, out)
    return out


def pad(x, pad, value=0.0):
    """cpu/pool.cpp:41-51: the planes inside a border of value (in x's type)."""
    pt, pl, pb, pr = pad
    return np.pad(x, [(0, 0)] * (x.ndim - 2) + [(pt, pb), (pl, pr)], constant_values=x.dtype.type(value))


def _values(x2, th):
    """wta.cpp:44-64 on one sample [T, L]: min and max over the sample,
    diff = (max - min) * 2, each neuron's value level + (x - min) / diff at its
    earliest step above th, else Limits<T>::min(); all in T."""
    dt = x2.dtype.type
    mn, mx = x2.min(), x2.max()
    diff = dt(dt(mx - mn) * dt(2))
    temp = np.full(x2.shape[1], lowest(dt), dt)
    t = x2.shape[0]
    for i in range(t - 1, -1, -1):
        row = x2[i]
        hit = row > dt(th)
        temp[hit] = dt(t - i - 1) + dt(dt(row[hit] - mn) / diff)
    return temp


def _stop(v):
    d = v - lowest(type(v))
    return abs(d) < _EPSILON


def convwta(x, radius, count, th=0.0):
    """wta.cpp:128-182 on [N, T, C, H, W]: per sample, k rounds of the first
    maximum; stop when it is within Epsilon of Limits<T>::min(); the box
    around the winner in every channel and the winner's channel are set to
    Limits<T>::min().  Returns (winners [N, count], counts [N])."""
    ry, rx = (radius, radius) if np.isscalar(radius) else radius
    n, t, c, h, w = x.shape
    dt = x.dtype.type
    win = np.zeros((n, count), WINNER)
    cnt = np.zeros(n, np.int64)
    for s in range(n):
        temp = _values(x[s].reshape(t, -1), th).reshape(c, h, w)
        k = 0
        while k < count:
            idx = int(temp.argmax())
            if _stop(temp.flat[idx]):
                break
            z, y, xx = idx // (h * w), (idx % (h * w)) // w, idx % w
            win[s, k] = (0, 0, z, y, xx)
            temp[:, max(0, y - ry):min(h, y + ry + 1), max(0, xx - rx):min(w, xx + rx + 1)] = lowest(dt)
            temp[z] = lowest(dt)
            k += 1
        cnt[s] = k
    return win, cnt


def fcwta(x, radius, count, th=0.0):
    """wta.cpp:87-126 on [N, T, F]."""
    n, t, f = x.shape
    dt = x.dtype.type
    win = np.zeros((n, count), WINNER)
    cnt = np.zeros(n, np.int64)
    for s in range(n):
        temp = _values(x[s], th)
        k = 0
        while k < count:
            idx = int(temp.argmax())
            if _stop(temp[idx]):
                break
            win[s, k] = (0, 0, 0, 0, idx)
            temp[max(0, idx - radius):min(f, idx + radius + 1)] = lowest(dt)
            k += 1
        cnt[s] = k
    return win, cnt


def _fma(a, b, c):
    """a * b + c rounded once to F64."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _config(c):
    """(positive, negative, stabilize, lower, upper) from a tuple as the C
    oracle takes it (defaults True, 0, 1) or an object with those fields."""
    if isinstance(c, (tuple, list)):
        return (float(c[0]), float(c[1]), bool(c[2]) if len(c) > 2 else True, float(c[3]) if len(c) > 3 else 0.0,
                float(c[4]) if len(c) > 4 else 1.0)
    return float(c.positive), float(c.negative), bool(c.stabilize), float(c.lower), float(c.upper)


def stdp_update(w, potentiate, cfg):
    """stdp.cpp:121-137: F64 arithmetic (`old + rate * value` contracted into
    one fused multiply-add by the reference's build), the clamp, the result
    rounded once to the weight's type."""
    positive, negative, stabilize, lower, upper = _config(cfg)
    dt = type(w)
    old = float(w)
    value = (old - lower) * (upper - old) if stabilize else 1.0
    value = _fma(positive if potentiate else negative, value, old)
    hi = value if value < upper else upper
    r = hi if lower < hi else lower
    return dt(np.float64(r))


def convstdp(spk_in, kernel, spk_out, configs, winners, counts, stride=1, pad=(0, 0, 0, 0)):
    """stdp.cpp:99-169 on U8 spikes with a kernel of any real type: for each
    sample's winners in order, the first input step at each tap of the
    winner's window (T where none, or in the pad), the first output step of
    the winner, then every weight of the winner's channel updated.  Returns
    the updated kernel (a copy)."""
    k = np.array(kernel, copy=True)
    n, t, cin, h, w = spk_in.shape
    cout, _, kh, kw = k.shape
    sy, sx = (stride, stride) if np.isscalar(stride) else stride
    pt, pl = pad[0], pad[1]
    nz_in, nz_out = spk_in != 0, spk_out != 0
    first_in = np.where(nz_in.any(axis=1), nz_in.argmax(axis=1), t)          # [n, cin, h, w]
    first_out = np.where(nz_out.any(axis=1), nz_out.argmax(axis=1), t)       # [n, cout, ho, wo]
    flat = k.reshape(cout, -1)
    for s in range(n):
        for j in range(int(counts[s])):
            wn = winners[s][j]
            if wn["c"] < 0:
                continue
            cfg = configs[int(wn["c"])]
            z, y, x = int(wn["z"]), int(wn["y"]), int(wn["x"])
            isum = np.full((cin, kh, kw), t, np.int64)
            for a in range(kh):
                for b in range(kw):
                    iy, ix = y * sy + a - pt, x * sx + b - pl
                    if 0 <= iy < h and 0 <= ix < w:
                        isum[:, a, b] = first_in[s, :, iy, ix]
            osum = int(first_out[s, z, y, x])
            row = flat[z]
            pot = (isum.reshape(-1) <= osum)
            for i in range(row.size):
                row[i] = stdp_update(row[i], bool(pot[i]), cfg)
    return k


def conv(orc, x, kernel, stride=1, pad=(0, 0, 0, 0)):
    """The plain conv with a kernel of any real type (opers.cpp:408-421: the
    input converted to the kernel's type, the output in the kernel's type) in
    the backend's fixed order (one fused multiply-add per tap in (cin, ky, kx)
    order from +0): F32 by the C oracle; F16 as that FP32 sum of the F16
    input and weights, rounded once to F16; F64 by the C oracle's F64 chain.
    x: [N, Cin, H, W]."""
    kt = kernel.dtype.type
    if kt == np.float32:
        return orc.conv(x, kernel, stride, pad)
    if kt == np.float16:
        xs = x if x.dtype == np.uint8 else x.astype(np.float16).astype(np.float32)
        return orc.conv(xs, kernel.astype(np.float32), stride, pad).astype(np.float16)
    return orc.conv_f64k(x, kernel, stride, pad)
