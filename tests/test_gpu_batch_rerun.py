"""Replayed batch executions (GPU, through the C ABI).

fastpls_hip_batch_rerun executes every fit of a batch again on the current
contents of its device inputs; the first call captures the whole launch
sequence into a HIP graph and every call replays it. A permutation test
writes new row permutations into the permutation buffers and reruns. Checked,
for precision 64 and 32:

* after each of several replays, with new permutations written in place
  and eager work (a fresh batch, synchronising copies) between the replays,
  every fit equals, bitwise, the fit of a freshly created batch with the same
  permutations (same kernels, same launch order, so the same arithmetic);
* a replay with the original permutations gives back the creation's results
  bitwise (the replay starts from cleared status words and recomputes every
  statistic, cross-product and Gram);
* label fits whose labels are permuted replay as well, and completed
  component counts and flags are reported for the replayed execution.
"""
import numpy as np
import pytest

from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu


def _dev(fp, a):
    return fp.DeviceBuffer.from_array(np.asfortranarray(a))


def _exports(batch, fits):
    return [{k: batch.export(i, k) for k in ("R", "Q", "T", "mX")} for i in range(len(fits))]


@pytest.mark.parametrize("precision", [64, 32])
def test_rerun_equals_fresh_batch(hip, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    rng = np.random.default_rng(21)
    keep, groups, fits, perms = [], [], [], []

    def add_group(X):
        b = _dev(fp, X.astype(dt))
        keep.append(b)
        groups.append((b.ptr, X.shape[0], X.shape[1]))
        return len(groups) - 1

    # a shared-X group of permutation refits (matrix-core passes)
    X, Y = latent_data(1200, 260, 2, factors=6, noise=0.1, seed=31)
    g = add_group(X)
    yb = _dev(fp, Y.astype(dt)); keep.append(yb)
    for r in range(12):
        pb = _dev(fp, rng.permutation(1200).astype(np.int32)); keep.append(pb)
        perms.append((pb, 1200))
        fits.append(dict(group=g, q=2, components=3 + r % 3, seed=50 + r, y=yb.ptr, perm=pb.ptr))
    # an unpermuted fit on the same X
    fits.append(dict(group=g, q=2, components=5, seed=7, y=yb.ptr))
    # a permuted label fit
    XC, lab = class_data(900, 120, 5, latent=4, seed=32, spread=1.5)
    g = add_group(XC)
    lb = _dev(fp, lab.astype(np.int32)); keep.append(lb)
    pb = _dev(fp, rng.permutation(900).astype(np.int32)); keep.append(pb)
    perms.append((pb, 900))
    fits.append(dict(group=g, q=5, components=4, seed=9, labels=lb.ptr, perm=pb.ptr))
    # an uneven singleton
    XS, YS = latent_data(700, 1100, 3, factors=5, noise=0.2, seed=33)
    g = add_group(XS)
    ys = _dev(fp, YS.astype(dt)); keep.append(ys)
    fits.append(dict(group=g, q=3, components=6, seed=11, y=ys.ptr))

    kw = dict(scaling=2, oversample=4, power=1)
    batch = fp.SimplsBatch(precision, groups, fits, **kw)
    done0, flags0, _ = batch.status()
    assert all(f == 0 for f in flags0), flags0
    first = _exports(batch, fits)
    originals = [pb.download((n,), np.int32) for pb, n in perms]

    for rnd in range(3):
        for pb, n in perms:
            pb.upload(rng.permutation(n).astype(np.int32))
        batch.rerun()
        done, flags, seconds = batch.status()
        assert all(f == 0 for f in flags), (rnd, flags)
        assert done == done0 and seconds > 0
        got = _exports(batch, fits)
        # eager work between replays: a fresh batch on the same permutations
        fresh = fp.SimplsBatch(precision, groups, fits, **kw)
        want = _exports(fresh, fits)
        fresh.close()
        for i in range(len(fits)):
            for k in ("R", "Q", "T", "mX"):
                np.testing.assert_array_equal(got[i][k], want[i][k], err_msg=f"{rnd} {i} {k}")
        # the permuted fits did change
        assert not np.array_equal(got[0]["R"], first[0]["R"])

    for (pb, n), orig in zip(perms, originals):
        pb.upload(orig)
    batch.rerun()
    again = _exports(batch, fits)
    for i in range(len(fits)):
        for k in ("R", "Q", "T", "mX"):
            np.testing.assert_array_equal(again[i][k], first[i][k], err_msg=f"orig {i} {k}")
    batch.close()
