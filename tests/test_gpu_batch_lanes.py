"""Batches split into lanes (GPU, through the C ABI).

A batch deals its shared-X groups to lanes, each run on its own stream of a
set that shares no hardware queue. Each fit's kernels and split plan depend on
its own group only, so a fit's arithmetic, and its results, must be those of
a one-lane batch. Checked, for precision 64 and 32, on 36 fits over five
groups (a permutation group, label fits, several response counts):

* 2 and 3 lanes (FASTPLS_HIP_BATCH_LANES) give every field of every fit,
  the completed counts and flags bitwise equal to one lane's;
* a graph replay of a 2-lane batch gives back its creation's results bitwise;
* the default is 2 lanes for 32 fits or more over two groups and 1 below 32
  fits; a lane count outside 1..4 is refused with a message.
"""
import os

import numpy as np
import pytest

from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

FIELDS = ("R", "Q", "T", "mX", "vX", "mY")


class _lanes:
    def __init__(self, value):
        self.value = value

    def __enter__(self):
        if self.value is not None:
            os.environ["FASTPLS_HIP_BATCH_LANES"] = str(self.value)

    def __exit__(self, *exc):
        os.environ.pop("FASTPLS_HIP_BATCH_LANES", None)


def _stage(fp, dt):
    rng = np.random.default_rng(41)
    keep, groups, fits = [], [], []

    def dev(a):
        b = fp.DeviceBuffer.from_array(np.asfortranarray(a))
        keep.append(b)
        return b

    def group(X):
        groups.append((dev(X.astype(dt)).ptr, X.shape[0], X.shape[1]))
        return len(groups) - 1

    X, Y = latent_data(1500, 300, 2, factors=6, noise=0.1, seed=42)
    g = group(X)
    yb = dev(Y.astype(dt))
    for r in range(14):
        pb = dev(rng.permutation(1500).astype(np.int32))
        fits.append(dict(group=g, q=2, components=2 + r % 4, seed=60 + r, y=yb.ptr, perm=pb.ptr))
    for d, (n, p, q) in enumerate([(900, 700, 1), (2500, 150, 5), (600, 900, 3)]):
        X, Y = latent_data(n, p, q, factors=5, noise=0.2, seed=43 + d)
        g = group(X)
        for r in range(5):
            fits.append(dict(group=g, q=q, components=2 + (r + d) % 5, seed=70 + 7 * d + r,
                             y=dev(Y[:, :q].astype(dt) * (1 + r)).ptr))
    XC, lab = class_data(1100, 200, 4, latent=4, seed=47, spread=1.5)
    g = group(XC)
    lb = dev(lab.astype(np.int32))
    for r in range(7):
        fits.append(dict(group=g, q=4, components=2 + r % 3, seed=90 + r, labels=lb.ptr))
    assert len(fits) == 36
    return keep, groups, fits


def _results(batch, n):
    done, flags, _ = batch.status()
    return list(done), list(flags), [{k: batch.export(i, k) for k in FIELDS} for i in range(n)]


def _same(a, b):
    assert a[0] == b[0] and a[1] == b[1]
    for fa, fb in zip(a[2], b[2]):
        for k in FIELDS:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)


@pytest.mark.parametrize("precision", [64, 32])
def test_lanes_equal_one_lane(hip, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    keep, groups, fits = _stage(fp, dt)
    kw = dict(scaling=2, oversample=6, power=1)
    with _lanes(1):
        one = fp.SimplsBatch(precision, groups, fits, **kw)
    assert one.lanes() == 1
    ref = _results(one, len(fits))
    assert all(f == 0 for f in ref[1])
    one.close()
    for L in (2, 3):
        with _lanes(L):
            b = fp.SimplsBatch(precision, groups, fits, **kw)
        assert b.lanes() == L
        _same(_results(b, len(fits)), ref)
        if L == 2:
            b.rerun()
            _same(_results(b, len(fits)), ref)
            b.rerun()
            _same(_results(b, len(fits)), ref)
        b.close()
    for x in keep:
        x.free()


def test_lane_defaults_and_limits(hip):
    fp = hip[0]
    keep, groups, fits = _stage(fp, np.float64)
    kw = dict(scaling=2, oversample=6, power=1)
    b = fp.SimplsBatch(64, groups, fits, **kw)
    assert b.lanes() == 2
    b.close()
    small = [f for f in fits if f["group"] == 0][:10]
    b = fp.SimplsBatch(64, groups[:1], small, **kw)
    assert b.lanes() == 1
    b.close()
    for bad in (0, 9):
        with _lanes(bad), pytest.raises(fp.HipError, match="FASTPLS_HIP_BATCH_LANES"):
            fp.SimplsBatch(64, groups, fits, **kw)
    for x in keep:
        x.free()
