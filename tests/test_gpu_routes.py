"""SIMPLS routes of the single-fit engine (GPU, through the C ABI).

* The cached predictor cross-product route (src/r_api.cpp:1165-1176: A >= 20,
  p <= n, n >= 8p, p <= 512; simpls.hpp:521-532, 600-620): the loadings come
  from X_sᵀX_s formed once, |t|² = rᵀ X_sᵀX_s r, and no loading pass runs.
  controls() reports it. The fit is compared at every prefix with the
  oracle's fit, which takes the same route by the same rule (and with the
  oracle's uncached core, which is the same mathematics): B = R_k Q_kᵀ and the
  fitted values T_k Q_kᵀ to 1e-9 (precision 64) and 1e-5 (precision 32, the
  oracle fitting the same FP32 values). All 22 components lie within the
  data's 30 latent factors, so both routes are well conditioned.
* Training-score retention (src/cuda_resident_simpls.cuh:213-216): a
  classification fit with retain_scores = 0, A >= 8, n >= 8p and scores
  above 256 MiB drops its training scores after the fit. Exporting T then
  fails with a message; LDA classification and predictions equal those of
  the same fit with its scores retained, bitwise.
* The square-sketch refresh: when every refresh is as wide as the response
  count (q <= block + oversample, so the sketch spans ℝ^q), the directions
  are the top left singular vectors of S, taken from the eigen-decomposition
  of SᵀS without the sketch, in the one-work-group refresh and in the
  multi-launch refresh of wide fits (FASTPLS_HIP_WIDE_P moves the threshold).
  Each fit is compared at every prefix with the oracle, which draws the sketch
  and runs the power iterations and QRs (1e-9 at precision 64, 1e-5 at
  precision 32 on the same FP32 values), and a fit with another seed is equal
  bitwise (no Gaussian draw is used). The classification fit takes candidate
  blocks of 6 (n p K >= 5e8) with K = 8 classes.
"""
import os

import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}


@pytest.mark.parametrize("precision", [64, 32])
def test_cached_crossprod_route(hip, oracle, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(4000, 200, 3, factors=30, noise=0.1, seed=71)
    X, Y = X.astype(dt), Y.astype(dt)
    kw = dict(precision=precision, oversample=6, power=2, seed=4)
    m = fp.ResidentModel.fit(X, Y, components=22, **kw)
    assert m.controls()["predictor_crossprod_cache"] == 1
    R, Q, T = m.export("R"), m.export("Q"), m.export("T")
    o = oracle.simpls_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=22, oversample=6,
                          power=2, seed=4)
    pr = oracle.prepare(X.astype(np.float64), Y.astype(np.float64), scaling=2)
    u = oracle.simpls_core(pr["xs"], pr["S"], 22, oversample=6, power=2, seed=4)
    for k in range(1, 23):
        B = R[:, :k].astype(np.float64) @ Q[:, :k].astype(np.float64).T
        F = T[:, :k].astype(np.float64) @ Q[:, :k].astype(np.float64).T
        for ref in (o, u):
            assert rel(B, ref["R"][:, :k] @ ref["Q"][:, :k].T) < TOL[precision], (k,)
            assert rel(F, ref["T"][:, :k] @ ref["Q"][:, :k].T) < TOL[precision], (k,)
    m.close()
    # below 20 components the two-pass route runs
    m19 = fp.ResidentModel.fit(X, Y, components=19, **kw)
    assert m19.controls()["predictor_crossprod_cache"] == 0
    m19.close()


def test_classification_scores_not_retained(hip):
    fp = hip[0]
    n, p, K, A = 4_300_000, 12, 3, 8  # n·A·8 B = 275 MB > 256 MiB, n >= 8p
    X, lab = class_data(n, p, K, latent=4, seed=72, spread=1.5)
    kw = dict(labels=lab, classes=K, precision=64, components=A, oversample=4, power=1, seed=2)
    kept = fp.ResidentModel.fit(X, retain_scores=True, **kw)
    dropped = fp.ResidentModel.fit(X, retain_scores=False, **kw)
    assert kept.export("T").shape == (n, A)
    with pytest.raises(fp.HipError, match="not retained"):
        dropped.export("T")
    np.testing.assert_array_equal(dropped.export("R"), kept.export("R"))
    Xq = np.asfortranarray(X[:6000])
    np.testing.assert_array_equal(dropped.classify_path(Xq, [4, 8], lda=True, top=2),
                                  kept.classify_path(Xq, [4, 8], lda=True, top=2))
    np.testing.assert_array_equal(dropped.predict_path(Xq, [2, 8]), kept.predict_path(Xq, [2, 8]))
    kept.close()
    dropped.close()
    # retained scores at smaller sizes whatever the request (the reference's rule)
    Xs, ls = class_data(2000, 12, K, latent=4, seed=73, spread=1.5)
    small = fp.ResidentModel.fit(Xs, labels=ls, classes=K, precision=64, components=A,
                                 oversample=4, power=1, seed=2, retain_scores=False)
    assert small.export("T").shape == (2000, A)
    small.close()


def _fit_route(fp, X, Y=None, wide=None, **kw):
    os.environ["FASTPLS_HIP_WIDE_P"] = str(wide)
    try:
        if Y is None:
            return fp.ResidentModel.fit(X, **kw)
        return fp.ResidentModel.fit(X, Y, **kw)
    finally:
        os.environ.pop("FASTPLS_HIP_WIDE_P", None)


def _prefixes_match(m, ref, A, tol):
    R, Q, T = (m.export(k).astype(np.float64) for k in ("R", "Q", "T"))
    for k in range(1, A + 1):
        assert rel(R[:, :k] @ Q[:, :k].T, ref["R"][:, :k] @ ref["Q"][:, :k].T) < tol, k
        assert rel(T[:, :k] @ Q[:, :k].T, ref["T"][:, :k] @ ref["Q"][:, :k].T) < tol, k


@pytest.mark.parametrize("wide", [1, 10 ** 9])
@pytest.mark.parametrize("precision", [64, 32])
def test_square_sketch_refresh(hip, oracle, wide, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    # all 14 components within the 16 factors: past the factors the fit is
    # ill-conditioned (the oracle differs from itself by ~1e-7 between seeds)
    X, Y = latent_data(1500, 400, 5, factors=16, noise=0.1, seed=74)
    X, Y = X.astype(dt), Y.astype(dt)
    kw = dict(precision=precision, components=14, oversample=6, power=2)
    m = _fit_route(fp, X, Y, wide=wide, seed=5, **kw)
    o = oracle.simpls_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=14, oversample=6,
                          power=2, seed=5)
    _prefixes_match(m, o, 14, TOL[precision])
    other = _fit_route(fp, X, Y, wide=wide, seed=999, **kw)
    np.testing.assert_array_equal(other.export("R"), m.export("R"))
    m.close(); other.close()


@pytest.mark.parametrize("wide", [1, 10 ** 9])
def test_square_sketch_refresh_classification_blocks(hip, oracle, wide):
    fp = hip[0]
    X, lab = class_data(12000, 5300, 8, latent=7, seed=75, spread=1.5)
    kw = dict(labels=lab, classes=8, precision=64, components=6, oversample=4, power=1)
    m = _fit_route(fp, X, wide=wide, seed=7, **kw)
    assert m.controls()["block"] == 6
    o = oracle.simpls_fit(X, labels=lab, classes=8, ncomp=6, oversample=4, power=1, seed=7)
    _prefixes_match(m, o, 6, TOL[64])
    m.close()
