"""The matrix-free (implicit cross-covariance) route on the GPU, through the C
ABI, against the FP64 oracle's restatement of simpls.hpp:768-1047 and
plssvd.hpp:222-248 (itself pinned to the compiled reference core in
tests/test_oracle.py).

Routing (R/main.R:3454-3463): no cross-covariance is formed once p·q·8 bytes
exceed 32 MiB; above 512 MiB of entries SIMPLS refreshes one direction at a
time by the dominant-direction power iteration (src/r_api.cpp:1480-1493).
Tolerances: precision 64, 1e-9 relative on B and Ŷ (1e-8 on the rank-one
route, whose unconverged power iteration amplifies rounding); precision 32
(FP32 X, bf16x6 products, FP64 small math), 1e-5 against the oracle on the
same FP32 values.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import latent_data

pytestmark = pytest.mark.gpu


def _coef(R, Q, k):
    return R[:, :k] @ Q[:, :k].T


def test_routing_threshold(hip):
    fp = hip[0]
    for q, implicit in ((209, 0), (210, 1)):
        X, Y = latent_data(60, 20000, q, factors=4, noise=0.1, seed=q)
        m = fp.ResidentModel.fit(X, Y, components=2, oversample=2, power=1, seed=1)
        assert m.controls()["implicit_operator"] == implicit, q
        m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_simpls_implicit_route(hip, oracle, precision):
    fp = hip[0]
    X, Y = latent_data(300, 20000, 220, factors=10, noise=0.1, seed=41, decay=0.8)
    dt = np.float64 if precision == 64 else np.float32
    Xp, Yp = X.astype(dt), Y.astype(dt)
    m = fp.ResidentModel.fit(Xp, Yp, precision=precision, components=6, oversample=10, power=2,
                             seed=4)
    assert m.controls()["implicit_operator"] == 1
    o = oracle.simpls_operator_fit(Xp.astype(np.float64), Yp.astype(np.float64), ncomp=6,
                                   oversample=10, power=2, seed=4, elem=precision // 8)
    assert o["rank_one"] == 0 and o["completed"] == 6
    R = m.export("R").astype(np.float64)
    Q = m.export("Q").astype(np.float64)
    tol = 1e-9 if precision == 64 else 1e-5
    for k in (1, 3, 6):
        assert rel(_coef(R, Q, k), _coef(o["R"], o["Q"], k)) < tol, k
    pred = m.predict_path(Xp[:100], [6])[:, :, 0].astype(np.float64)
    Xs = (Xp[:100].astype(np.float64) - o["mx"]) / o["sx"]
    assert rel(pred, Xs @ _coef(o["R"], o["Q"], 6) + o["my"]) < tol * 10
    m.close()


def test_simpls_rank_one_route(hip, oracle):
    """p·q·8 = 544 MB > 512 MiB: block 1, the dominant-direction power
    iteration on the projected operator, two re-orthogonalisation passes."""
    fp = hip[0]
    X, Y = latent_data(200, 20000, 3400, factors=6, noise=0.1, seed=42, decay=0.6)
    m = fp.ResidentModel.fit(X, Y, components=4, oversample=4, power=3, seed=9)
    c = m.controls()
    assert c["implicit_operator"] == 1 and c["block"] == 1
    o = oracle.simpls_operator_fit(X, Y, ncomp=4, oversample=4, power=3, seed=9, elem=8)
    assert o["rank_one"] == 1 and o["completed"] == 4
    R = m.export("R")
    Q = m.export("Q")
    for k in (1, 2, 4):
        assert rel(_coef(R, Q, k), _coef(o["R"], o["Q"], k)) < 1e-8, k
    m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_plssvd_implicit_route(hip, oracle, precision):
    fp = hip[0]
    X, Y = latent_data(300, 20000, 220, factors=10, noise=0.1, seed=43, decay=0.8)
    dt = np.float64 if precision == 64 else np.float32
    Xp, Yp = X.astype(dt), Y.astype(dt)
    m = fp.ResidentModel.fit(Xp, Yp, method=fp.PLSSVD, precision=precision, components=5,
                             oversample=3, power=2, seed=7)
    assert m.controls()["implicit_operator"] == 1
    o = oracle.plssvd_operator_fit(Xp.astype(np.float64), Yp.astype(np.float64), ncomp=5,
                                   oversample=3, power=2, seed=7)
    tol = 1e-9 if precision == 64 else 1e-5
    pred = m.predict_path(Xp[:100], [2, 5]).astype(np.float64)
    Xs = (Xp[:100].astype(np.float64) - o["mx"]) / o["sx"]
    for j, k in enumerate((2, 5)):
        yh = Xs @ (o["R"][:, :k] @ o["W"][k - 1]) + o["my"]
        assert rel(pred[:, :, j], yh) < tol * 10, k
    m.close()
