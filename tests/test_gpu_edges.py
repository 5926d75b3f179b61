"""Reference edge rules on the GPU, through the C ABI.

* Scaling 1 (centring only, R's public default, R/main.R:9746-9748) and 3
  (none: X uncentred, Y centred; supervised.hpp:63-124,
  classification.hpp:86-184) for single SIMPLS and PLS-SVD fits (against the
  compiled reference core's goldens at precision 64, and the oracle on the
  same FP32 values at precision 32), for a grouped batch (shared-X
  permutation refits on both the VALU and the matrix-core width classes, and
  singletons) and for the CV drivers (against fold-by-fold oracle
  restatements).
* The LDA ridge ladder past 1e-8 and its exhaustion (lda.hpp:86-131) and the
  PLS-SVD prefix solve's pivoted fallback and failure (plssvd.hpp:138-164,
  linalg.hpp:57-177), driven exactly from given moments through the device
  kernels the fits use.
* Non-finite inputs (NaN in Y or X) end a fit with an error, never a
  silent result.

Tolerances: precision 64, 1e-8 against the goldens (1e-7 where the direction
depends on the Gaussian sketch) and 1e-9 against the oracle; precision 32,
1e-5 against the oracle; the small operators 1e-12 (1e-9 where the ridged
covariance has condition ~5e5) and the exact ridge rung.
"""
import numpy as np
import pytest

from conftest import align_columns, load_golden, rel
from oracle.oracle import class_data, latent_data
from test_gpu_cv import _oracle_cv_classification, _oracle_cv_regression

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}


def _coef(R, Q, k):
    return R[:, :k].astype(np.float64) @ Q[:, :k].astype(np.float64).T


# ------------------------------------------------------------ single fits
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["simpls_scaling1", "simpls_scaling3", "simpls_labels_scaling1"])
def test_simpls_scaling_single_fit(hip, oracle, name, precision):
    fp = hip[0]
    g = load_golden(name)
    dt = np.float64 if precision == 64 else np.float32
    sc = int(g["scaling"])
    kw = dict(components=int(g["ncomp"]), scaling=sc, oversample=int(g["oversample"]),
              power=int(g["power"]), seed=int(g["seed"]))
    okw = dict(ncomp=int(g["ncomp"]), scaling=sc, oversample=int(g["oversample"]),
               power=int(g["power"]), seed=int(g["seed"]))
    X = g["X"].astype(dt)
    if "labels" in g:
        m = fp.ResidentModel.fit(X, labels=g["labels"], classes=int(g["classes"]),
                                 precision=precision, **kw)
        o = oracle.simpls_fit(X.astype(np.float64), labels=g["labels"], classes=int(g["classes"]),
                              **okw)
    else:
        m = fp.ResidentModel.fit(X, g["Y"].astype(dt), precision=precision, **kw)
        o = oracle.simpls_fit(X.astype(np.float64), g["Y"].astype(dt).astype(np.float64), **okw)
    R, Q = m.export("R").astype(np.float64), m.export("Q").astype(np.float64)
    gtol = 1e-7 if "labels" in g else 1e-8
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(_coef(R, Q, k), _coef(o["R"], o["Q"], k)) < TOL[precision], (k, precision)
        if precision == 64:
            assert rel(_coef(R, Q, k), _coef(g["R"], g["Q"], k)) < gtol, k
    mx, sx = m.export("mX").astype(np.float64), m.export("vX").astype(np.float64)
    np.testing.assert_allclose(sx, 1.0, rtol=0)  # scaling 1 and 3 never scale
    if sc == 3:
        np.testing.assert_array_equal(mx, 0.0)
    else:
        np.testing.assert_allclose(mx, g["mx"], rtol=1e-6 if precision == 32 else 1e-12,
                                   atol=1e-12)
    # predictions follow the same centring rule
    yh = m.predict(X[:20], int(g["ncomp"])).astype(np.float64)
    Xs = (X[:20].astype(np.float64) - o["mx"]) / o["sx"]
    want = Xs @ _coef(o["R"], o["Q"], int(g["ncomp"])) + o["my"]
    assert rel(yh, want) < TOL[precision]
    m.close()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["plssvd_scaling1", "plssvd_scaling3"])
def test_plssvd_scaling_single_fit(hip, oracle, name, precision):
    fp = hip[0]
    g = load_golden(name)
    dt = np.float64 if precision == 64 else np.float32
    sc = int(g["scaling"])
    X, Y = g["X"].astype(dt), g["Y"].astype(dt)
    m = fp.ResidentModel.fit(X, Y, method=fp.PLSSVD, precision=precision,
                             components=int(g["ncomp"]), scaling=sc,
                             oversample=int(g["oversample"]), power=int(g["power"]),
                             seed=int(g["seed"]))
    o = oracle.plssvd_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=int(g["ncomp"]),
                          scaling=sc, oversample=int(g["oversample"]), power=int(g["power"]),
                          seed=int(g["seed"]))
    R = m.export("R").astype(np.float64)
    ref = g if precision == 64 else o
    assert rel(align_columns(R, ref["R"]), ref["R"]) < (1e-7 if precision == 64 else 1e-5)
    Xq = X[:30]
    Xs = (Xq.astype(np.float64) - o["mx"]) / o["sx"]
    for k in range(1, int(g["ncomp"]) + 1):
        yh = m.predict(Xq, k).astype(np.float64)
        want = Xs @ (o["R"][:, :k] @ o["W"][k - 1]) + o["my"]
        assert rel(yh, want) < (1e-8 if precision == 64 else 1e-5), k
    np.testing.assert_allclose(m.export("vX").astype(np.float64), 1.0, rtol=0)
    m.close()


# ------------------------------------------------------------------ batch
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("scaling", [1, 3])
def test_batch_scaling_matches_oracle(hip, oracle, scaling, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    rng = np.random.default_rng(70 + scaling)
    keep, groups, fits, ref = [], [], [], []

    def group(X):
        b = fp.DeviceBuffer.from_array(np.asfortranarray(X.astype(dt)))
        keep.append(b)
        groups.append((b.ptr, X.shape[0], X.shape[1]))
        return len(groups) - 1

    # shared X with 3 refits (VALU width) and with 10 refits (matrix-core width)
    for n, p, q, A, refits in [(500, 90, 1, 5, 3), (640, 150, 1, 4, 10)]:
        X, Y = latent_data(n, p, q, factors=8, noise=0.1, seed=80 + refits)
        g = group(X)
        yb = fp.DeviceBuffer.from_array(np.asfortranarray(Y.astype(dt))); keep.append(yb)
        for r in range(refits):
            f = dict(group=g, q=q, components=A, seed=300 + len(fits), y=yb.ptr)
            Yu = Y
            if r:
                perm = rng.permutation(n).astype(np.int32)
                pb = fp.DeviceBuffer.from_array(perm); keep.append(pb)
                f["perm"] = pb.ptr
                Yu = Y[perm]
            fits.append(f)
            ref.append((X, Yu, A, f["seed"]))
    for n, p, q, A in [(300, 420, 2, 5), (900, 60, 4, 6)]:
        X, Y = latent_data(n, p, q, factors=8, noise=0.1, seed=n + p)
        g = group(X)
        yb = fp.DeviceBuffer.from_array(np.asfortranarray(Y.astype(dt))); keep.append(yb)
        fits.append(dict(group=g, q=q, components=A, seed=300 + len(fits), y=yb.ptr))
        ref.append((X, Y, A, fits[-1]["seed"]))
    b = fp.SimplsBatch(precision, groups, fits, scaling=scaling, oversample=6, power=2)
    done, flags, _ = b.status()
    assert all(fl == 0 for fl in flags), flags
    for i, (X, Y, A, seed) in enumerate(ref):
        o = oracle.simpls_fit(X.astype(dt).astype(np.float64), Y.astype(dt).astype(np.float64),
                              ncomp=A, scaling=scaling, oversample=6, power=2, seed=seed)
        R, Q, T = (b.export(i, f, f64=True) for f in ("R", "Q", "T"))
        for k in range(1, A + 1):
            assert rel(_coef(R, Q, k), _coef(o["R"], o["Q"], k)) < TOL[precision], (i, k)
            assert rel(_coef(T, Q, k), _coef(o["T"], o["Q"], k)) < TOL[precision], (i, k)
        mx = b.export(i, "mX", f64=True)
        if scaling == 3:
            np.testing.assert_array_equal(mx, 0.0)
        np.testing.assert_array_equal(b.export(i, "vX", f64=True), 1.0)
    b.close()


# ------------------------------------------------------------------- CV
@pytest.mark.parametrize("scaling", [1, 3])
@pytest.mark.parametrize("method", [0, 1])
def test_cv_regression_scaling(hip, oracle, method, scaling):
    fp = hip[0]
    X, Y = latent_data(240, 30, 3, factors=4, noise=0.3, seed=90 + scaling)
    folds = (np.arange(240) % 4 + 1).astype(np.int32)
    np.random.default_rng(9).shuffle(folds)
    prefixes = [1, 3]
    run = fp.simpls_cv_regression if method == 0 else fp.plssvd_cv_regression
    got = run(X, Y, folds, prefixes, precision=64, scaling=scaling, metric=3, oversample=4,
              power=2, seed=7)
    want = _oracle_cv_regression(oracle, method, X, Y, folds, prefixes, seed=7, oversample=4,
                                 power=2, scaling=scaling)
    np.testing.assert_array_equal(got["status"], want["status"])
    for key in ("q2", "rmsd", "observed_r2"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-9)
    assert rel(got["predictions"], want["predictions"]) < 1e-9


@pytest.mark.parametrize("scaling", [1, 3])
def test_cv_classification_scaling(hip, oracle, scaling):
    fp = hip[0]
    X, lab = class_data(400, 40, 4, latent=5, seed=95, spread=0.8)
    folds = (np.arange(400) % 4 + 1).astype(np.int32)
    np.random.default_rng(10).shuffle(folds)
    prefixes = [1, 3]
    got = fp.simpls_cv_classification(X, lab, folds, 4, prefixes, precision=64, scaling=scaling,
                                      lda=True, oversample=2, power=2, seed=4, store_scores=True)
    want = _oracle_cv_classification(oracle, 0, X, lab, folds, 4, prefixes, True, seed=4,
                                     oversample=2, power=2, scaling=scaling)
    np.testing.assert_array_equal(got["status"], want["status"])
    np.testing.assert_array_equal(got["predictions"], want["predictions"])
    assert rel(got["scores"], want["scores"]) < 1e-9


# ------------------------------------------------------- small operators
def test_lda_ridge_ladder_on_device(hip, oracle):
    fp = hip[0]
    g = load_golden("lda_ladder")
    for k in (1, 2, 3):
        got = fp.lda_moments(g["gram"], g["sums"], g["counts"], int(g["n"]), k)
        assert got["ridge"] == pytest.approx(float(g[f"ridge{k}"]), rel=1e-12), k
        # at prefix 3 the ridged pooled covariance has condition ~5e5 (its
        # smallest eigenvalue is 4e-6 of a trace of 3): rounding of 1e-16
        # reaches the weights as ~5e-11 relative
        tol = 1e-12 if k < 3 else 1e-9
        np.testing.assert_allclose(got["linear"], g[f"linear{k}"], rtol=tol, atol=1e-12)
        np.testing.assert_allclose(got["constants"], g[f"constants{k}"], rtol=tol, atol=1e-12)
    # the 1e-5 rung: relative ridge = ridge / (trace(pooled) / 3)
    o = oracle.lda_moments(g["gram"], g["sums"], g["counts"], int(g["n"]), 3)
    assert o["relative"] == 1e-5
    with pytest.raises(fp.HipError, match="every regularization level"):
        fp.lda_moments(g["gram_fail"], g["sums"], g["counts"], int(g["n"]), 3)
    with pytest.raises(fp.HipError, match="do not sum to n"):
        fp.lda_moments(g["gram"], g["sums"], g["counts"], int(g["n"]) + 1, 3)


def test_plssvd_pivoted_prefix_solve_on_device(hip):
    fp = hip[0]
    g = load_golden("plssvd_pivoted")
    assert list(g["route"]) == [0, 1, 1, 1]  # prefixes 2-4 need the pivoted solve
    W = fp.plssvd_weights(g["G"], g["sigma"], g["Q"])
    for k in range(1, 5):
        np.testing.assert_allclose(W[k - 1], g[f"W{k}"], rtol=1e-12, atol=1e-12)
    with pytest.raises(fp.HipError, match="latent solve failed"):
        fp.plssvd_weights(g["G_singular"], g["sigma"][:2], g["Q"][:, :2])


# ------------------------------------------------------------- non-finite
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("where", ["Y", "X"])
def test_nonfinite_inputs_are_errors(hip, method, where):
    fp = hip[0]
    X, Y = latent_data(200, 30, 3, factors=4, seed=99)
    if where == "Y":
        Y[17, 1] = np.nan
    else:
        X[5, 3] = np.nan
    with pytest.raises(fp.HipError, match="non-finite|failed"):
        fp.ResidentModel.fit(X, Y, method=method, precision=64, components=2, scaling=2,
                             oversample=4, power=1, seed=1)


def test_nonfinite_response_flags_batched_fit(hip):
    fp = hip[0]
    X, Y = latent_data(200, 30, 2, factors=4, seed=98)
    Ybad = Y.copy(); Ybad[3, 0] = np.nan
    xb = fp.DeviceBuffer.from_array(X)
    yg, yb = fp.DeviceBuffer.from_array(Y), fp.DeviceBuffer.from_array(Ybad)
    b = fp.SimplsBatch(64, [(xb.ptr, 200, 30)],
                       [dict(group=0, q=2, components=2, seed=1, y=yg.ptr),
                        dict(group=0, q=2, components=2, seed=1, y=yb.ptr)],
                       scaling=2, oversample=4, power=1)
    done, flags, _ = b.status()
    assert flags[0] == 0 and done[0] == 2
    assert flags[1] != 0  # the NaN fit reports itself; its neighbour is unaffected
    b.close()
