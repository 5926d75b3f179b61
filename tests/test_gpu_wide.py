"""The multi-launch refresh and update of wide single SIMPLS fits (GPU, through
the C ABI).

Fits with p >= 1536 run refresh_directions (simpls.hpp:226-368) and the
component update (simpls.hpp:600-756) as several launches whose products use
the whole device, instead of one work-group per fit. FASTPLS_HIP_WIDE_P
replaces the threshold, so the route is checked here on small problems too:
regression with the right Gram (q <= p), regression without it (q > p), and
classification with candidate blocks (n p K >= 5e8). Each is compared at
every prefix with the FP64 oracle (B = R_k Q_kᵀ and T_k Q_kᵀ; 1e-9 at
precision 64, 1e-5 at precision 32 on the same FP32 values) and with the
one-work-group route (1e-9 / 1e-6: the same arithmetic in other summation
orders). A p = 5000 fit takes the route at the default threshold.
"""
import os

import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}
SAME = {64: 1e-9, 32: 1e-6}


def _fit(fp, X, Y=None, labels=None, K=None, wide=None, **kw):
    if wide is not None:
        os.environ["FASTPLS_HIP_WIDE_P"] = str(wide)
    try:
        if labels is not None:
            return fp.ResidentModel.fit(X, labels=labels, classes=K, **kw)
        return fp.ResidentModel.fit(X, Y, **kw)
    finally:
        os.environ.pop("FASTPLS_HIP_WIDE_P", None)


def _check(m, ref, A, tol):
    R, Q, T = (m.export(k).astype(np.float64) for k in ("R", "Q", "T"))
    for k in range(1, A + 1):
        assert rel(R[:, :k] @ Q[:, :k].T, ref["R"][:, :k] @ ref["Q"][:, :k].T) < tol, k
        assert rel(T[:, :k] @ Q[:, :k].T, ref["T"][:, :k] @ ref["Q"][:, :k].T) < tol, k


CASES = {
    "right_gram": dict(n=900, p=300, q=4, A=8),
    "no_right_gram": dict(n=500, p=30, q=40, A=6),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", [64, 32])
def test_wide_route_regression(hip, oracle, case, precision):
    fp = hip[0]
    c = CASES[case]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(c["n"], c["p"], c["q"], factors=10, noise=0.1, seed=81)
    X, Y = X.astype(dt), Y.astype(dt)
    kw = dict(precision=precision, components=c["A"], oversample=5, power=2, seed=6)
    wide = _fit(fp, X, Y, wide=1, **kw)
    narrow = _fit(fp, X, Y, wide=10 ** 9, **kw)
    o = oracle.simpls_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=c["A"], oversample=5,
                          power=2, seed=6)
    _check(wide, o, c["A"], TOL[precision])
    _check(narrow, o, c["A"], TOL[precision])
    _check(wide, {k: narrow.export(k).astype(np.float64) for k in ("R", "Q", "T")}, c["A"],
           SAME[precision])
    wide.close(); narrow.close()


def test_wide_route_classification_blocks(hip, oracle):
    fp = hip[0]
    X, lab = class_data(6000, 1500, 60, latent=8, seed=82, spread=1.0)
    kw = dict(precision=64, components=8, oversample=4, power=1, seed=7)
    wide = _fit(fp, X, labels=lab, K=60, wide=1, **kw)
    assert wide.controls()["block"] > 1
    o = oracle.simpls_fit(X, labels=lab, classes=60, ncomp=8, oversample=4, power=1, seed=7)
    _check(wide, o, 8, TOL[64])
    Xq = X[:500]
    narrow = _fit(fp, X, labels=lab, K=60, wide=10 ** 9, **kw)
    np.testing.assert_array_equal(wide.classify_path(Xq, [4, 8], lda=True),
                                  narrow.classify_path(Xq, [4, 8], lda=True))
    wide.close(); narrow.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_wide_route_default_threshold(hip, oracle, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(3000, 5000, 4, factors=12, noise=0.1, seed=83)
    X, Y = X.astype(dt), Y.astype(dt)
    m = fp.ResidentModel.fit(X, Y, precision=precision, components=10, oversample=6, power=2,
                             seed=8)
    o = oracle.simpls_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=10, oversample=6,
                          power=2, seed=8)
    _check(m, o, 10, TOL[precision])
    m.close()
