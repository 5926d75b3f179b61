"""Grouped multi-fit execution and the batched CV driver (GPU, through the C ABI).

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
* fastpls_hip_simpls_cv_regression must reproduce a fold-by-fold
  restatement on the FP64 oracle (src/cuda_resident_api.cuh:772-1023: the
  same fold rows, the shared seed, SSE / training-mean SS per fold, the
  metric ladder and the prediction layout).
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
        # column means: FP64 sums of per-lane partials (FP32 partials over 16
        # rows on the bf16x3 path), so different split layouts agree to
        # rounding relative to the column scale
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


def _oracle_cv(oracle, X, Y, folds, prefixes, seed, oversample, power, scaling=2):
    """Fold-by-fold restatement of src/cuda_resident_api.cuh:772-1023."""
    n, q = Y.shape
    F = folds.max()
    sse = np.zeros(len(prefixes)); tss = np.zeros(len(prefixes)); cnt = np.zeros(len(prefixes))
    pred = np.zeros((n, q, len(prefixes)))
    status = np.zeros(F, dtype=np.int32)
    for f in range(F):
        te = np.where(folds == f + 1)[0]
        if te.size == 0:
            status[f] = 2
            continue
        tr = np.where(folds != f + 1)[0]
        m = oracle.simpls_fit(X[tr], Y[tr], ncomp=max(prefixes), scaling=scaling,
                              oversample=oversample, power=power, seed=seed)
        Xs = (X[te] - m["mx"]) / m["sx"]
        for j, k in enumerate(prefixes):
            yh = Xs @ (m["R"][:, :k] @ m["Q"][:, :k].T) + m["my"]
            pred[te, :, j] = yh
            sse[j] += ((Y[te] - yh) ** 2).sum()
            tss[j] += ((Y[te] - m["my"]) ** 2).sum()
            cnt[j] += te.size * q
        status[f] = 1
    total = ((Y - Y.mean(0)) ** 2).sum()
    return dict(status=status, sse=sse, q2=1 - sse / tss, rmsd=np.sqrt(sse / cnt),
                observed_r2=1 - sse / total, predictions=pred)


@pytest.mark.parametrize("metric", [2, 3, 4])
def test_cv_regression_matches_oracle(hip, oracle, metric):
    fp = hip[0]
    X, Y = latent_data(300, 40, 2, factors=5, noise=0.3, seed=6)
    folds = (np.arange(300) % 5 + 1).astype(np.int32)
    np.random.default_rng(2).shuffle(folds)
    prefixes = [1, 3, 5]
    got = fp.simpls_cv_regression(X, Y, folds, prefixes, precision=64, metric=metric,
                                  oversample=6, power=2, seed=5)
    want = _oracle_cv(oracle, X, Y, folds, prefixes, seed=5, oversample=6, power=2)
    np.testing.assert_array_equal(got["status"], want["status"])
    np.testing.assert_allclose(got["q2"], want["q2"], rtol=1e-9)
    np.testing.assert_allclose(got["rmsd"], want["rmsd"], rtol=1e-9)
    np.testing.assert_allclose(got["observed_r2"], want["observed_r2"], rtol=1e-9)
    expect = {2: want["observed_r2"], 3: want["q2"], 4: want["rmsd"]}[metric]
    np.testing.assert_allclose(got["metrics"], expect, rtol=1e-9)
    assert rel(got["predictions"], want["predictions"]) < 1e-9


def test_cv_regression_empty_fold_status(hip, oracle):
    fp = hip[0]
    X, Y = latent_data(120, 15, 1, factors=3, seed=7)
    folds = np.where(np.arange(120) % 2 == 0, 1, 3).astype(np.int32)  # fold 2 empty
    got = fp.simpls_cv_regression(X, Y, folds, [2], precision=64, metric=3, oversample=4,
                                  power=1, seed=1)
    assert list(got["status"]) == [1, 2, 1]
    with pytest.raises(fp.HipError, match="strictly increasing"):
        fp.simpls_cv_regression(X, Y, folds, [2, 2], precision=64)
