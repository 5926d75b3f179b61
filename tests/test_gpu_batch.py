"""Grouped multi-fit execution (GPU, through the C ABI).

* A batch of uneven fits must give, for every fit, the same model as fitting
  it alone through fastpls_hip_simpls_create: groups that share X (row
  permutations of the responses, wide enough for the matrix-core passes),
  label fits with block refresh and batch geometry, and singletons.
  Tolerance: 1e-9 (precision 64) and 1e-5 (precision 32) on the fitted
  values T_k Q_kᵀ at every prefix, and on B = R_k Q_kᵀ at every prefix in
  FP64 and, in FP32, at prefixes within the latent rank of the data for fits
  of unpermuted responses. Past the rank, and everywhere for a row
  permutation of the responses (which removes the X-Y relation: the
  permutation refits are pure-noise fits), components have tiny singular
  gaps, so FP32-level differences between the batched and the single-fit
  arithmetic (FP32-accumulated matrix-core passes against FP64 VALU passes)
  move them by ~1e-5..1e-4 relative while the fitted values agree.

The batched CV driver is tested with the other CV entries in test_gpu_cv.py.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}


def coef(R, Q, k):
    return R[:, :k].astype(np.float64) @ Q[:, :k].astype(np.float64).T


def _dev(fp, a):
    return fp.DeviceBuffer.from_array(np.asfortranarray(a))


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_equals_alone(hip, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    rng = np.random.default_rng(11)
    keep = []  # device buffers alive for the batch's lifetime

    groups, fits, alone = [], [], []

    def add_group(X):
        b = _dev(fp, X.astype(dt))
        keep.append(b)
        groups.append((b.ptr, X.shape[0], X.shape[1]))
        return len(groups) - 1

    # group 0: three uneven regression fits on one X, one with a row permutation
    XA, YA = latent_data(600, 80, 3, factors=6, noise=0.2, seed=1)
    g = add_group(XA)
    for q, A, seed, permute in [(2, 5, 3, False), (3, 7, 4, True), (1, 3, 5, False)]:
        Y = YA[:, :q]
        perm = rng.permutation(600).astype(np.int32) if permute else None
        yb = _dev(fp, Y.astype(dt)); keep.append(yb)
        f = dict(group=g, q=q, components=A, seed=seed, y=yb.ptr)
        if perm is not None:
            pb = _dev(fp, perm); keep.append(pb)
            f["perm"] = pb.ptr
        fits.append(f)
        alone.append(dict(X=XA, Y=Y[perm] if perm is not None else Y, A=A, seed=seed,
                          rank=0 if permute else 6))

    # group 1: ten row permutations of one response (fourteen right-hand
    # columns: the per-step passes run on the matrix cores)
    XD, YD = latent_data(1500, 300, 2, factors=8, noise=0.1, seed=2)
    g = add_group(XD)
    yb = _dev(fp, YD.astype(dt)); keep.append(yb)
    for r in range(10):
        perm = rng.permutation(1500).astype(np.int32)
        pb = _dev(fp, perm); keep.append(pb)
        A = 4 if r < 7 else 6
        fits.append(dict(group=g, q=2, components=A, seed=100 + r, y=yb.ptr, perm=pb.ptr))
        alone.append(dict(X=XD, Y=YD[perm], A=A, seed=100 + r, rank=0))

    # group 2: classification with candidate blocks (n p K >= 5e8)
    XC, lab = class_data(6000, 1500, 60, latent=8, seed=3, spread=1.0)
    g = add_group(XC)
    lb = _dev(fp, lab.astype(np.int32)); keep.append(lb)
    fits.append(dict(group=g, q=60, components=8, seed=7, labels=lb.ptr))
    alone.append(dict(X=XC, labels=lab, K=60, A=8, seed=7, rank=8))

    # group 3: a small label singleton
    XS, labs = class_data(400, 50, 4, latent=3, seed=4, spread=2.0)
    g = add_group(XS)
    lb = _dev(fp, labs.astype(np.int32)); keep.append(lb)
    fits.append(dict(group=g, q=4, components=3, seed=9, labels=lb.ptr))
    alone.append(dict(X=XS, labels=labs, K=4, A=3, seed=9, rank=3))

    batch = fp.SimplsBatch(precision, groups, fits, scaling=2, oversample=4, power=1)
    done, flags, seconds = batch.status()
    assert all(f == 0 for f in flags), flags
    assert seconds > 0
    for i, a in enumerate(alone):
        assert done[i] == a["A"], (i, done[i])
        if "labels" in a:
            m = fp.ResidentModel.fit(a["X"].astype(dt), labels=a["labels"], classes=a["K"],
                                     precision=precision, components=a["A"], scaling=2,
                                     oversample=4, power=1, seed=a["seed"])
        else:
            m = fp.ResidentModel.fit(a["X"].astype(dt), a["Y"].astype(dt), precision=precision,
                                     components=a["A"], scaling=2, oversample=4, power=1,
                                     seed=a["seed"])
        Rb, Qb, Tb = batch.export(i, "R"), batch.export(i, "Q"), batch.export(i, "T")
        Ra, Qa, Ta = m.export("R"), m.export("Q"), m.export("T")
        for k in range(1, a["A"] + 1):
            assert rel(coef(Tb, Qb, k), coef(Ta, Qa, k)) < TOL[precision], (precision, i, k)
            if precision == 64 or k <= a["rank"]:
                assert rel(coef(Rb, Qb, k), coef(Ra, Qa, k)) < TOL[precision], (precision, i, k)
        # column means: FP64 sums of per-lane partials (FP32 partials over
        # one 64-row step on the matrix-core path), so different split
        # layouts agree to rounding relative to the column scale
        scale = float(np.mean(m.export("vX")))
        np.testing.assert_allclose(batch.export(i, "mX"), m.export("mX"), rtol=0,
                                   atol=(1e-12 if precision == 64 else 1e-6) * scale)
        m.close()
    batch.close()


def test_batch_rejects_bad_descriptions(hip):
    fp = hip[0]
    X, Y = latent_data(50, 10, 1, seed=5)
    xb = _dev(fp, X); yb = _dev(fp, Y)
    with pytest.raises(fp.HipError, match="group"):
        fp.SimplsBatch(64, [(xb.ptr, 50, 10)], [dict(group=3, q=1, components=2, y=yb.ptr)])
    with pytest.raises(fp.HipError, match="dimensions"):
        fp.SimplsBatch(64, [(xb.ptr, 50, 10)], [dict(group=0, q=1, components=11, y=yb.ptr)])
