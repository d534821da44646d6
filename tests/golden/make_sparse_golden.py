"""Writes tests/golden/sparse_vectors.npz from the reference's own sparse code.

Every expected output comes from oracle/_ref/libspyker_ref.so, compiled from
the reference's sources (src/spyker/sparse/*.cpp, cpu/wta.cpp) by
oracle/Makefile in this container.  Event lists are stored as the oracle
holds them: off[n*t + 1] and events packed z << 16 | y << 8 | x.

Run:  make -C oracle && python tests/golden/make_sparse_golden.py
Inputs are seeded (PCG64); the file is small enough to commit.
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from spyker_amd import synthetic  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sparse_vectors.npz")


def put(v, name, sp):
    v[name + "_off"], v[name + "_ev"], v[name + "_shape"] = sp.off, sp.ev, np.array(sp.shape, np.int64)


def main():
    R = oracle.Reference()
    rng = np.random.Generator(np.random.PCG64(2025))
    v = {}

    # code: bar images with zeros and a tie block, sorted and latency
    img = synthetic.bars(3, 20, seed=7)[:, None].repeat(2, axis=1)
    img[:, 1] *= 0.5
    img[img < 0.2] = 0
    img[0, 0, :4, :4] = 0.75  # ties
    v["code_in"] = img
    put(v, "code_sort", R.sparse_code(img, 12, True))
    put(v, "code_latency", R.sparse_code(img, 12, False))

    # dense -> sparse -> dense / gather on DVS-like frames with values 0..3
    ev = synthetic.events(3, 10, 24, seed=8, density=0.05) * rng.integers(1, 4, (3, 10, 2, 24, 24)).astype(np.uint8)
    v["dense_in"] = ev
    put(v, "from_dense_0", R.sparse_from_dense(ev, 0))
    put(v, "from_dense_1", R.sparse_from_dense(ev, 1))
    sp = R.sparse_from_dense(ev, 0)
    v["to_dense"] = R.sparse_to_dense(sp)
    v["gather"] = R.sparse_gather(sp)

    # pad, pool (list order of the output is not channel-major), inhibit, convwta
    put(v, "pad", R.sparse_pad(sp, (2, 1, 2, 1)))
    put(v, "pool2_pad10", R.sparse_pool(sp, 2, (1, 0, 0, 0)))
    put(v, "pool3", R.sparse_pool(sp, 3))
    put(v, "inhibit", R.sparse_inhibit(sp))
    w, c = R.sparse_convwta(sp, 1, 5)
    v["convwta_r1_k5_win"], v["convwta_r1_k5_cnt"] = w, c
    w, c = R.sparse_convwta(sp, (0, 2), 3)
    v["convwta_r02_k3_win"], v["convwta_r02_k3_cnt"] = w, c

    # conv 2->8 5x5 pad 2 on sorted lists, on pooled (unsorted) lists, and
    # with a negative threshold (everything fires at the first step with events)
    k = np.clip(rng.normal(0.8, 0.05, (8, 2, 5, 5)), 0, 1).astype(np.float32)
    v["conv_kernel"] = k
    out = R.sparse_conv(sp, k, 6.0, (2, 2, 2, 2))
    put(v, "conv", out)
    pooled = R.sparse_pool(sp, 2)
    put(v, "conv_pooled", R.sparse_conv(pooled, k, 4.0, (2, 2, 2, 2)))
    put(v, "conv_negative", R.sparse_conv(sp, k[:3] - 0.9, -0.5, (1, 0, 2, 3)))

    # STDP on the conv's winners, R-STDP configs (one unstabilised with wide
    # bounds) and a skipped winner
    w, c = R.sparse_convwta(out, 2, 4)
    w = w.copy()
    w["c"][:, 1::2] = 1
    w["c"][0, 0] = -1
    cfg = [(0.004, -0.003), (0.01, -0.002, False, -1.0, 2.0)]
    v["stdp_win"], v["stdp_cnt"] = w, c
    v["stdp_kernel"] = R.sparse_stdp(sp, k, cfg, w, c, (2, 2, 2, 2))

    # dense fcwta on integer spikes (the densified sparse output)
    d = R.sparse_to_dense(out).reshape(3, 10, -1)
    for r, cnt in ((0, 16), (3, 5)):
        w, c = R.fcwta(d, r, cnt)
        v["fcwta_u8_r%d_k%d_win" % (r, cnt)], v["fcwta_u8_r%d_k%d_cnt" % (r, cnt)] = w, c

    np.savez_compressed(OUT, **v)
    print("wrote", OUT, sum(a.nbytes for a in v.values()), "bytes")


if __name__ == "__main__":
    main()
