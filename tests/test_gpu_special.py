"""Resident OPLS and nonlinear kernel PLS on the GPU, through the C ABI.

Checked against the reference core's own compositions (fixtures written by
tests/golden/make_golden.py from fit_opls / predict_opls, the label filter
and fit_kernelpls / predict_kernelpls) and against the FP64 oracle, which
restates them and is pinned to those fixtures in tests/test_oracle.py.

Tolerances (relative Frobenius error): precision 64 1e-9 on B = R_k Q_kᵀ
and on predictions; precision 32 (inputs FP32, filter and kernel in FP64,
the inner feature matrix stored FP32) 1e-5 against the oracle run on the
same FP32 values; labels identical except rows whose top-2 margin is below
1e-6.
"""
import numpy as np
import pytest

from conftest import align_columns, load_golden, rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}


def as_prec(a, precision):
    return np.asfortranarray(np.asarray(a, dtype=np.float32 if precision == 32 else np.float64))


def hip_coef(model, k):
    R = model.export("R").astype(np.float64)
    Q = model.export("Q").astype(np.float64)
    return R[:, :k] @ Q[:, :k].T


def filter_rows(m, x):
    """Standardised rows filtered by the oracle's orthogonal components
    (opls.hpp:467-509)."""
    xs = (np.asarray(x, dtype=np.float64) - m["mx"]) / m["sx"]
    for c in range(m["W"].shape[1]):
        xs = xs - np.outer(xs @ m["W"][:, c], m["P"][:, c])
    return xs


def opls_predict(m, x, k):
    return filter_rows(m, x) @ m["R"][:, :k] @ m["Q"][:, :k].T + m["my"]


# --------------------------------------------------------------------- OPLS
@pytest.mark.parametrize("name", ["opls_dense", "opls_sketch"])
def test_opls_matches_reference_core_fixture(hip, name):
    fp = hip[0]
    g = load_golden(name)
    k = int(g["prefix"])
    with fp.ResidentModel.fit_opls(g["X"], g["Y"], precision=64, components=int(g["ncomp"]),
                                   north=int(g["north"]), scaling=int(g["scaling"]),
                                   oversample=int(g["oversample"]), power=int(g["power"]),
                                   seed=int(g["seed"])) as m:
        print(name, "OPLS", g["X"].shape, flush=True)
        pred = m.predict_path(g["Xt"], [k])[:, :, 0]
        assert rel(pred, g["pred"]) < 1e-9
        np.testing.assert_allclose(m.export("mX"), g["mx"], rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(m.export("vX"), g["sx"], rtol=1e-12)
        # training scores of the inner model, up to a sign per component
        T = m.export("T")
        assert rel(align_columns(T, g["T"]), g["T"]) < 1e-8


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("north,scaling", [(0, 2), (1, 1), (2, 2)])
def test_opls_regression_matches_oracle(hip, oracle, precision, north, scaling):
    fp = hip[0]
    X, Y = latent_data(300, 40, 3, factors=10, seed=41)
    Xt, _ = latent_data(50, 40, 3, factors=10, seed=42)
    X, Y, Xt = as_prec(X, precision), as_prec(Y, precision), as_prec(Xt, precision)
    A = 4
    kw = dict(ncomp=A, north=north, scaling=scaling, oversample=4, power=2, seed=9)
    o = oracle.opls_fit(X.astype(np.float64), Y.astype(np.float64), **kw)
    print("OPLS", precision, north, scaling, flush=True)
    with fp.ResidentModel.fit_opls(X, Y, precision=precision, components=A, north=north,
                                   scaling=scaling, oversample=4, power=2, seed=9) as m:
        tol = TOL[precision]
        pred = m.predict_path(Xt, list(range(1, A + 1)))
        for k in range(1, A + 1):
            # the exported weights are R_eff: standardised rows -> scores
            assert rel(hip_coef(m, k), o["Reff"][:, :k] @ o["Q"][:, :k].T) < tol, k
            assert rel(pred[:, :, k - 1], opls_predict(o, Xt, k)) < tol, k
        # exports come back in the handle's precision
        etol = 1e-12 if precision == 64 else 2e-7
        np.testing.assert_allclose(m.export("mX"), o["mx"], rtol=etol, atol=etol)
        np.testing.assert_allclose(m.export("vX"), o["sx"], rtol=etol)
        np.testing.assert_allclose(m.export("mY"), o["my"], rtol=etol, atol=etol)
        # project() gives the inner scores of the filtered rows
        sc = m.project(Xt, A)
        ref_sc = filter_rows(o, Xt) @ o["R"]
        assert rel(align_columns(sc, ref_sc), ref_sc) < tol
        # response sums of the inner model on filtered rows
        sums = m.response_sums(Xt, A, y=Y[:50])
        yh = opls_predict(o, Xt, A)
        yo = Y[:50].astype(np.float64)
        sse = ((yo - yh) ** 2).sum(axis=0)
        assert rel(sums[0], sse) < max(tol, 1e-9)
        # variance of the inner model: loadings of the filtered training rows
        Xf = filter_rows(o, X)
        var = oracle.variance(Xf, o["T"])
        P = m.export("P")
        assert rel(align_columns(P, var["P"]), var["P"]) < max(tol, 1e-8)


@pytest.mark.parametrize("precision", [64, 32])
def test_opls_labels_match_oracle(hip, oracle, precision):
    fp = hip[0]
    g = load_golden("opls_labels")
    X = as_prec(g["X"], precision)
    lab = g["labels"]
    kw = dict(ncomp=int(g["ncomp"]), north=int(g["north"]), oversample=int(g["oversample"]),
              power=int(g["power"]), seed=int(g["seed"]))
    o = oracle.opls_fit(X.astype(np.float64), labels=lab, classes=4, **kw)
    if precision == 64:
        # the oracle restates the reference core's label filter + SIMPLS
        assert rel(align_columns(o["W"], g["W"]), g["W"]) < 1e-10
    print("OPLS labels", precision, flush=True)
    with fp.ResidentModel.fit_opls(X, labels=lab, classes=4, precision=precision,
                                   components=kw["ncomp"], north=kw["north"],
                                   oversample=kw["oversample"], power=kw["power"],
                                   seed=kw["seed"]) as m:
        A = kw["ncomp"]
        tol = TOL[precision]
        for k in range(1, A + 1):
            assert rel(hip_coef(m, k), o["Reff"][:, :k] @ o["Q"][:, :k].T) < tol, k
        labels, resp = m.classify_response_path(X, [A], top=2)
        ref = opls_predict(o, X.astype(np.float64), A)
        assert rel(resp[:, :, 0], ref) < tol
        srt = np.sort(ref, axis=1)
        margin = srt[:, -1] - srt[:, -2]
        want = np.argmax(ref, axis=1) + 1
        ok = (labels[:, 0, 0] == want) | (margin < 1e-6)
        assert ok.all()
        # LDA on the inner model's training scores
        lda_labels = m.classify_path(X, [A], lda=True)[:, 0, 0]
        scores = filter_rows(o, X.astype(np.float64)) @ o["R"]
        lda = oracle.lda(o["T"], lab, 4, A, scores)
        d = np.sort(lda["disc"], axis=1)
        near = (d[:, -1] - d[:, -2]) < 1e-6
        assert ((lda_labels == lda["labels"]) | near).all()


def test_opls_compact_keeps_queries(hip, oracle):
    fp = hip[0]
    X, Y = latent_data(200, 30, 2, factors=6, seed=51)
    o = oracle.opls_fit(X, Y, ncomp=3, north=1, oversample=4, power=2, seed=3)
    with fp.ResidentModel.fit_opls(X, Y, precision=64, components=3, north=1, oversample=4,
                                   power=2, seed=3) as m:
        m.compact()
        assert rel(m.predict(X[:20], 3), opls_predict(o, X[:20], 3)) < 1e-9
        with pytest.raises(fp.HipError):
            m.export("P")


# --------------------------------------------------------------- kernel PLS
@pytest.mark.parametrize("name", ["kernelpls_rbf", "kernelpls_poly"])
def test_kernelpls_matches_reference_core_fixture(hip, name):
    fp = hip[0]
    g = load_golden(name)
    k = int(g["prefix"])
    print(name, flush=True)
    with fp.ResidentModel.fit_kernelpls(g["X"], g["Y"], precision=64, components=int(g["ncomp"]),
                                        kernel=int(g["kernel"]), gamma=float(g["gamma"]),
                                        degree=int(g["degree"]), coef0=float(g["coef0"]),
                                        scaling=int(g["scaling"]),
                                        oversample=int(g["oversample"]), power=int(g["power"]),
                                        seed=int(g["seed"])) as m:
        pred = m.predict_path(g["Xt"], [k])[:, :, 0]
        assert rel(pred, g["pred"]) < 1e-9
        for kk in range(1, int(g["ncomp"]) + 1):
            B = g["R"][:, :kk] @ g["Q"][:, :kk].T
            assert rel(hip_coef(m, kk), B) < 1e-9, kk
        np.testing.assert_allclose(m.export("mX"), g["mx"], rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(m.export("vX"), g["sx"], rtol=1e-12)
        assert m.export("R").shape == (g["X"].shape[0], int(g["ncomp"]))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("kernel,gamma,degree", [(2, 0.02, 3), (3, 0.05, 2)])
def test_kernelpls_matches_oracle(hip, oracle, precision, kernel, gamma, degree):
    fp = hip[0]
    X, Y = latent_data(400, 20, 2, factors=6, seed=61)
    Xt, _ = latent_data(60, 20, 2, factors=6, seed=62)
    X, Y, Xt = as_prec(X, precision), as_prec(Y, precision), as_prec(Xt, precision)
    A = 4
    kw = dict(ncomp=A, kernel=kernel, gamma=gamma, degree=degree, coef0=1.0, oversample=4,
              power=2, seed=13)
    print("kernel PLS", precision, kernel, flush=True)
    with fp.ResidentModel.fit_kernelpls(X, Y, precision=precision, components=A, kernel=kernel,
                                        gamma=gamma, degree=degree, coef0=1.0, oversample=4,
                                        power=2, seed=13) as m:
        tol = TOL[precision]
        pred = m.predict_path(Xt, list(range(1, A + 1)))
        for k in range(1, A + 1):
            o = oracle.kernelpls_fit(X.astype(np.float64), Y.astype(np.float64), xtest=Xt, prefix=k,
                                     **kw)
            assert rel(pred[:, :, k - 1], o["pred"]) < tol, k
            assert rel(hip_coef(m, k), o["R"][:, :k] @ o["Q"][:, :k].T) < tol, k
        T = m.export("T")
        assert rel(align_columns(T, o["T"]), o["T"]) < max(tol, 1e-8)


def test_kernelpls_labels_match_oracle(hip, oracle):
    fp = hip[0]
    X, lab = class_data(240, 15, 3, latent=4, seed=71)
    o = oracle.kernelpls_fit(X, labels=lab, classes=3, ncomp=3, kernel=2, gamma=0.05,
                             oversample=1, power=1, seed=4, xtest=X, prefix=3)
    print("kernel PLS labels", flush=True)
    with fp.ResidentModel.fit_kernelpls(X, labels=lab, classes=3, precision=64, components=3,
                                        kernel=2, gamma=0.05, oversample=1, power=1,
                                        seed=4) as m:
        labels, resp = m.classify_response_path(X, [3])
        assert rel(resp[:, :, 0], o["pred"]) < 1e-9
        srt = np.sort(o["pred"], axis=1)
        near = (srt[:, -1] - srt[:, -2]) < 1e-6
        assert ((labels[:, 0, 0] == np.argmax(o["pred"], axis=1) + 1) | near).all()


def test_special_handle_errors(hip, monkeypatch):
    fp = hip[0]
    X, Y = latent_data(50, 8, 2, factors=3, seed=81)
    with pytest.raises(fp.HipError, match="OPLS"):
        fp.ResidentModel.fit_opls(X, Y, components=2, north=-1)
    with pytest.raises(fp.HipError, match="kernel"):
        fp.ResidentModel.fit_kernelpls(X, Y, components=2, kernel=4)
    with pytest.raises(fp.HipError, match="kernel"):
        fp.ResidentModel.fit_kernelpls(X, Y, components=2, kernel=2, gamma=0.0)
    with pytest.raises(fp.HipError, match="kernel"):
        fp.ResidentModel.fit_kernelpls(X, Y, components=50, kernel=2)
    with fp.ResidentModel.fit_kernelpls(X, Y, components=2, kernel=2, gamma=0.1) as m:
        with pytest.raises(fp.HipError):
            m.predict_path(X, [3])  # prefix beyond the components
        # the test-kernel workspace guard (src/cuda_resident_special.cuh:388-398)
        monkeypatch.setenv("FASTPLS_HIP_QUERY_FREE_BYTES", "1000")
        with pytest.raises(fp.HipError, match="guarded HIP workspace"):
            m.predict(X, 2)
